// Filters for unorganized point clouds (a merged multi-view cloud): SOR on
// the exact k nearest neighbours, MLS and Gaussian smoothing on the
// neighbours within a radius. Not in the reference, which filters organized
// clouds only (JSLAB_MLS_Filter.cpp, JSLAB_PhaseEpipolarReconstructor_V2.cpp);
// the per-point arithmetic is the organized filters'.
//
// Grid: cells of edge `voxel` (psgc_voxel.h); the points are sorted by cell
// key with a stable radix sort (ties keep index order), the occupied cells
// found by run-length encoding. A workgroup takes one occupied cell: the
// block of cells around it (3x3x3; 5x5x5 for radius filters on half-radius
// cells) is looked up once (binary search over the sorted keys), its points
// streamed through LDS in chunks (one point per thread), and every thread scans each
// chunk for one point of the centre cell. Candidates come in a fixed order
// (cells dz, dy, dx = -R..R, points in grid order), the order of the CPU
// restatement.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "psgc_device.h"
#include "psgc_internal.h"
#include "psgc_math.h"
#include "psgc_tiled.h"
#include "psgc_voxel.h"

namespace psgc {
namespace {

constexpr int CB = 256;    // threads per workgroup

struct CellTab {
  const uint64_t* ukey;
  const int32_t* start;
  const int32_t* count;
  int32_t ncells;
};

__device__ inline int32_t cell_find(const CellTab& t, uint64_t key) {
  int32_t lo = 0, hi = t.ncells;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (ld_global(t.ukey, mid) < key) lo = mid + 1;
    else hi = mid;
  }
  return (lo < t.ncells && ld_global(t.ukey, lo) == key) ? lo : -1;
}

__device__ inline bool point_ok(const float* xyz, const uint8_t* valid, int64_t i) {
  return (!valid || ld_global(valid, i) != 0) &&
         voxel_point_ok(ld_global(xyz, i * 3), ld_global(xyz, i * 3 + 1), ld_global(xyz, i * 3 + 2));
}

// ---------------------------------------------------------------- grid ----

// Per-workgroup bounding box and count of the participating points.
__global__ void __launch_bounds__(CB) cloud_bbox_kernel(const float* __restrict__ xyz,
                                                        const uint8_t* __restrict__ valid,
                                                        int64_t n, float* __restrict__ part) {
  __shared__ float red[7][CB];
  float v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.f};
  for (int64_t i = (int64_t)blockIdx.x * CB + threadIdx.x; i < n; i += (int64_t)gridDim.x * CB) {
    if (!point_ok(xyz, valid, i)) continue;
    for (int c = 0; c < 3; ++c) {
      const float x = ld_global(xyz, i * 3 + c);
      v[c] = fminf(v[c], x);
      v[3 + c] = fmaxf(v[3 + c], x);
    }
    v[6] += 1.f;
  }
  for (int c = 0; c < 7; ++c) red[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int h = CB / 2; h >= 1; h /= 2) {
    if (threadIdx.x < h)
      for (int c = 0; c < 7; ++c) {
        const float a = red[c][threadIdx.x], b = red[c][threadIdx.x + h];
        red[c][threadIdx.x] = c < 3 ? fminf(a, b) : (c < 6 ? fmaxf(a, b) : a + b);
      }
    __syncthreads();
  }
  if (threadIdx.x < 7) st_global(part, (size_t)blockIdx.x * 8 + threadIdx.x, red[threadIdx.x][0]);
}

// Cell key per point (participating ones; the others get `sentinel`, which
// sorts after every cell) and the identity permutation.
__global__ void cloud_keys_kernel(const float* __restrict__ xyz, const uint8_t* __restrict__ valid,
                                  int64_t n, VoxelFrame f, uint64_t sentinel,
                                  uint64_t* __restrict__ key, int32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = sentinel;
  if (point_ok(xyz, valid, i))
    k = voxel_key_of(f, ld_global(xyz, i * 3), ld_global(xyz, i * 3 + 1), ld_global(xyz, i * 3 + 2));
  st_global(key, i, k);
  st_global(idx, i, (int32_t)i);
}

// Sorted points with their index in the fourth lane.
__global__ void cloud_gather_kernel(const float* __restrict__ xyz, const int32_t* __restrict__ idx,
                                    int32_t np, float4* __restrict__ pts) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np) return;
  const int32_t i = ld_global(idx, j);
  pts[j] = make_float4(ld_global(xyz, (int64_t)i * 3), ld_global(xyz, (int64_t)i * 3 + 1),
                       ld_global(xyz, (int64_t)i * 3 + 2), __int_as_float(i));
}

// ------------------------------------------------------- block staging ----

// The (2R+1)^3 cells around a cell: first point and running counts.
template <int R>
struct BlockCells {
  static constexpr int W = 2 * R + 1, N = W * W * W;
  int32_t st[N], pre[N + 1];
};

// The 27 cells ordered by distance class from the centre: itself, the 6
// face neighbours, the 12 edge neighbours, the 8 corners (index dz*9 + dy*3
// + dx with offsets shifted by one).
__constant__ int8_t kNearFirst[27] = {13, 4, 10, 12, 14, 16, 22, 1, 3, 5, 7, 9, 11,
                                      15, 17, 19, 21, 23, 25, 0, 2, 6, 8, 18, 20, 24, 26};

// The cells around `key` (absent ones empty) and the running counts, in the
// canonical order (dz, dy, dx = -R..R), or (R = 1) nearest class first.
// TB >= 64 threads; the first wave scans the counts, two cells per lane.
template <int R, int TB, bool NEAR_FIRST = false>
__device__ inline void load_block(const CellTab& t, const VoxelFrame& f, uint64_t key,
                                  BlockCells<R>* b) {
  using B = BlockCells<R>;
  static_assert(B::N <= 128 && TB >= 64 && (!NEAR_FIRST || R == 1), "block shape");
  __shared__ int32_t cnt[128];
  for (int slot = threadIdx.x; slot < 128; slot += TB) {
    int32_t c = -1;
    if (slot < B::N) {
      const int d = NEAR_FIRST ? kNearFirst[slot] : slot;
      const int64_t off = (d % B::W - R) +
                          f.X * ((d / B::W) % B::W - R + f.Y * (int64_t)(d / (B::W * B::W) - R));
      c = cell_find(t, (uint64_t)((int64_t)key + off));
      b->st[slot] = c >= 0 ? ld_global(t.start, c) : 0;
    }
    cnt[slot] = c >= 0 ? ld_global(t.count, c) : 0;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const int l = threadIdx.x;
    const int32_t a0 = cnt[2 * l], a1 = cnt[2 * l + 1];
    int32_t inc = a0 + a1;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t v = __shfl_up(inc, o, 64);
      if (l >= o) inc += v;
    }
    const int32_t ex = inc - a0 - a1;
    if (2 * l <= B::N) b->pre[2 * l] = ex;
    if (2 * l + 1 <= B::N) b->pre[2 * l + 1] = ex + a0;
  }
  __syncthreads();
}

// Candidates [cb, cb + len) of the block into tile (cooperative, TB threads).
template <int R, int TB>
__device__ inline void stage_chunk(const float4* __restrict__ pts, const BlockCells<R>* b,
                                   int32_t cb, int32_t len, float4* tile) {
  for (int32_t j = threadIdx.x; j < len; j += TB) {
    const int32_t g = cb + j;
    int lo = 0, hi = BlockCells<R>::N - 1;  // last cell with pre <= g
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (b->pre[mid] <= g) lo = mid;
      else hi = mid - 1;
    }
    tile[j] = pts[b->st[lo] + (g - b->pre[lo])];
  }
}

// ------------------------------------------------------------------ kNN ----

// K > 0: exactly K nearest kept in ascending order (a value replaces the
// largest and bubbles down); K == 0: k <= 64 at run time.
template <int K>
struct Knn {
  static constexpr int N = K > 0 ? K : 64;
  float best[N];
  int cnt;
  __device__ void init() {
#pragma unroll
    for (int i = 0; i < N; ++i) best[i] = INFINITY;
    cnt = 0;
  }
  __device__ void add(float v, int k) {
    if (K > 0) {
      ++cnt;
      if (!(v < best[N - 1])) return;
      best[N - 1] = v;
#pragma unroll
      for (int t = N - 1; t > 0; --t) {
        const float a = best[t - 1], b = best[t];
        if (!(b < a)) break;
        best[t - 1] = b;
        best[t] = a;
      }
    } else {
      const int have = cnt < k ? cnt : k;
      ++cnt;
      if (have == k && !(v < best[k - 1])) return;
#pragma unroll
      for (int t = 0; t < N; ++t) {
        if (t < have) {
          const float b = best[t];
          const bool sw = v < b;
          best[t] = sw ? v : b;
          v = sw ? b : v;
        } else if (t == have) {
          best[t] = v;
        }
      }
    }
  }
  // The k-th smallest when there are k candidates, else +inf.
  __device__ float kth(int k) const {
    if (cnt < k) return INFINITY;
    float r = best[0];
#pragma unroll
    for (int t = 0; t < N; ++t)
      if (t == k - 1) r = best[t];
    return r;
  }
  __device__ double mean_dist(int k) const {
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < N; ++t)
      if (t < k) s += sqrt_of_float(best[t]);
    return s / k;
  }
};

// The NB smallest of a candidate stream (NB a power of two, 32 or 64),
// ascending, taken in groups of G (16 or 32) with fixed compare-exchange
// networks and no data-dependent branches: a group is sorted (bitonic sort),
// its reverse paired with the list's top G (min of each pair: with the list
// ascending and the group descending the result is a bitonic sequence
// holding the NB smallest of both), and the list re-sorted by a bitonic
// merge. The first K
// <= NB entries are then exactly the K smallest seen, as a one-by-one
// insertion keeps them (the same multiset, so the same ascending list), at a
// cost that does not grow with how far each value travels. NaN candidates
// enter as +inf.
template <int NB>
struct KnnNet {
  static_assert(NB == 32 || NB == 64, "list of 32 or 64");
  float best[NB];
  __device__ void init() {
#pragma unroll
    for (int i = 0; i < NB; ++i) best[i] = INFINITY;
  }
  __device__ static void cx(float& a, float& b) {
    const float lo = fminf(a, b), hi = fmaxf(a, b);
    a = lo;
    b = hi;
  }
  template <int G>
  __device__ __attribute__((always_inline)) void merge(float (&g)[G]) {
    static_assert(G == 16 || G == 32, "groups of 16 or 32");
#pragma unroll
    for (int size = 2; size <= G; size <<= 1)
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
        for (int i = 0; i < G; ++i) {
          const int j = i ^ stride;
          if (j > i) {
            if ((i & size) == 0) cx(g[i], g[j]);
            else cx(g[j], g[i]);
          }
        }
#pragma unroll
    for (int i = 0; i < G; ++i) best[NB - G + i] = fminf(best[NB - G + i], g[G - 1 - i]);
#pragma unroll
    for (int stride = NB >> 1; stride > 0; stride >>= 1)
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int j = i ^ stride;
        if (j > i) cx(best[i], best[j]);
      }
  }
  __device__ double mean_dist(int k) const {
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < NB; ++t)
      if (t < k) s += sqrt_of_float(best[t]);
    return s / k;
  }
};

// First pass: the 3x3x3 block. Points it cannot settle (fewer than k
// candidates, or a k-th distance beyond the point's bound: one voxel plus
// its distance to the nearest face of its cell) are flagged in `again` at
// their grid position, with the block's k-th distance in `kth_of`.
// TB threads per workgroup (one cell at a time), chunks of chunk_of(TB)
// candidates (one per thread: the smaller LDS tile keeps more workgroups
// resident; 2 or 4 per thread measured slower), candidate queues of QL per
// lane (32: half the merges of 16, whose networks dominate the pass).
constexpr int chunk_of(int tb) { return tb; }
constexpr int QL = 32;

template <int K, int TB>
__global__ void __launch_bounds__(TB, (K >= 50 ? 4 : 1)) cloud_knn_kernel(const float4* __restrict__ pts, CellTab t,
                                                       VoxelFrame f, int k, float lim,
                                                       float limmax, double* __restrict__ md,
                                                       uint8_t* __restrict__ again,
                                                       float* __restrict__ kth_of) {
  constexpr int CH = chunk_of(TB);
  __shared__ float4 tile[CH];
  __shared__ BlockCells<1> bc;
  __shared__ float queue[K > 0 ? QL : 1][TB];  // per-lane candidate queues
  const XcdRange xr = xcd_range(t.ncells);
  for (int64_t c = xr.first; c < xr.end; c += xr.stride) {
    // nearest cells first: the k-th distance falls early and later
    // candidates rarely enter the list (the selected multiset does not
    // depend on the order)
    load_block<1, TB, true>(t, f, ld_global(t.ukey, c), &bc);
    const int32_t M = bc.pre[BlockCells<1>::N];
    const int32_t c0 = ld_global(t.start, c), cn = ld_global(t.count, c);
    for (int32_t qb = 0; qb < cn; qb += TB) {
      const bool active = qb + (int32_t)threadIdx.x < cn;
      const int32_t q = c0 + qb + threadIdx.x;
      const float4 qp = active ? pts[q] : make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
      const int qi = __float_as_int(qp.w);
      // the point's own bound on the block search (at least the block's,
      // at most the one of the widest search the retry may make)
      const float limq = fminf(fmaxf(lim, cloud_knn_point_limit(f, qp.x, qp.y, qp.z)), limmax);
      if constexpr (K > 0) {
        // Each lane queues the candidates below its current K-th distance
        // (which only falls, so a candidate at or above it can never enter)
        // in QL LDS slots of its own; the queues are merged into the sorted
        // lists when one fills (uniform per wave), so the merge networks run
        // about once per QL accepted candidates instead of once per QL read.
        KnnNet<(K <= 32 ? 32 : 64)> sel;
        sel.init();
        int qn = 0;
        auto flush = [&]() __attribute__((always_inline)) {
          float g[QL];
#pragma unroll
          for (int u = 0; u < QL; ++u) g[u] = u < qn ? queue[u][threadIdx.x] : INFINITY;
          sel.merge(g);
          qn = 0;
        };
        for (int32_t cb = 0; cb < M; cb += CH) {
          const int32_t len = min(CH, M - cb);
          __syncthreads();  // the previous chunk's readers are done
          stage_chunk<1, TB>(pts, &bc, cb, len, tile);
          __syncthreads();
          // four candidates per step (their reads issue together), after
          // making room for four in every queue
          for (int32_t j0 = 0; j0 < len; j0 += 4) {
            if (__any(qn > QL - 4)) flush();
            float4 p[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) p[u] = tile[min(j0 + u, len - 1)];
            const float kth = sel.best[K - 1];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const float pk[3] = {p[u].x, p[u].y, p[u].z};
              const float e = dist2(qp.x, qp.y, qp.z, pk);
              if (active && j0 + u < len && __float_as_int(p[u].w) != qi && e < kth)
                queue[qn++][threadIdx.x] = e;
            }
          }
        }
        if (__any(qn > 0)) flush();
        if (active) {
          if (sel.best[K - 1] <= limq) {
            st_global(md, qi, sel.mean_dist(K));
          } else {
            st_global(again, q, (uint8_t)1);
            st_global(kth_of, q, sel.best[K - 1]);
          }
        }
      } else {
        Knn<K> sel;
        sel.init();
        for (int32_t cb = 0; cb < M; cb += CH) {
          const int32_t len = min(CH, M - cb);
          __syncthreads();  // the previous chunk's readers are done
          stage_chunk<1, TB>(pts, &bc, cb, len, tile);
          __syncthreads();
          if (active)
            for (int32_t j = 0; j < len; ++j) {
              const float4 p = tile[j];
              if (__float_as_int(p.w) == qi) continue;
              const float pk[3] = {p.x, p.y, p.z};
              sel.add(dist2(qp.x, qp.y, qp.z, pk), k);
            }
        }
        if (active) {
          if (sel.kth(k) <= limq) {
            st_global(md, qi, sel.mean_dist(k));
          } else {
            st_global(again, q, (uint8_t)1);
            st_global(kth_of, q, sel.kth(k));
          }
        }
      }
    }
    __syncthreads();  // block cells are reloaded for the next cell
  }
}

// First cell whose key is not below `key` (cells are sorted by key).
__device__ inline int32_t cell_lower_bound(const CellTab& t, uint64_t key) {
  int32_t lo = 0, hi = t.ncells;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (ld_global(t.ukey, mid) < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Cell coordinate of a double position (a search box bound; clamped to the
// frame).
__device__ inline int64_t box_cell(double v, float o, double inv, int64_t n) {
  const double c = floor((v - (double)o) * inv);
  return c < 0. ? 0 : (c > (double)(n - 1) ? n - 1 : (int64_t)c);
}

__device__ inline int wave_min_i(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = min(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ inline int wave_max_i(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m, 64));
  return v;
}

// Retry: for each listed point (grid positions in ascending order, so the
// points of one cell and of its x-neighbours are consecutive), the exact k
// nearest within the ball that must hold them: the k-th distance the 3x3x3
// block found bounds the true one from above (infinite when the block held
// fewer than k candidates), and a k-th nearest beyond rmax voxels makes the
// point not eligible anyway, so the radius is the smaller of the two,
// widened by 1e-5 (far above the float rounding of the distances and the
// cell assignment). Every point within it lies in the box of cells it spans.
// A wave takes 64 consecutive listed points; where their boxes lie close
// together it walks their union row by row (one key search per row of
// cells, then consecutive keys), every lane reading the same candidate at a
// time: one memory transaction per candidate for the wave, where boxes
// walked lane by lane take one per lane (as the other waves do). A lane
// keeps a candidate only up to its own bound (the block's k-th distance,
// which the k nearest cannot exceed, and the limit of rmax voxels), so the
// k it keeps are those of its own box either way. A k-th squared distance up to
// the limit settles the point, as the block search widening by shells up to
// rmax would (both find the true k nearest whenever it lies within that
// limit). Unsettled points keep md = -1 (not eligible).
template <int K>
__global__ void __launch_bounds__(CB) cloud_knn_retry_kernel(const float4* __restrict__ pts,
                                                             CellTab t, VoxelFrame f, int k,
                                                             int rmax,
                                                             const int32_t* __restrict__ retry,
                                                             const float* __restrict__ kth_of,
                                                             const int32_t* __restrict__ nretry,
                                                             double* __restrict__ md) {
  const int32_t nr = ld_global(nretry, 0);
  const float lim = cloud_knn_limit(f.voxel, rmax);
  const int lane = (int)threadIdx.x & 63;
  for (int64_t base = (int64_t)blockIdx.x * CB + ((int)threadIdx.x - lane); base < nr;
       base += (int64_t)gridDim.x * CB) {
    const int64_t r = base + lane;
    const bool act = r < nr;
    const int32_t q = act ? ld_global(retry, r) : ld_global(retry, base);
    const float4 qp = pts[q];
    const int qi = __float_as_int(qp.w);
    const float k1 = ld_global(kth_of, q);
    double rho = (double)f.voxel * rmax;
    if (k1 < INFINITY) rho = fmin(rho, sqrt((double)k1));
    rho *= 1.0 + 1e-5;
    // the lane's box of cells (coordinates below 2^30) and the wave's union
    const int x0 = (int)box_cell(qp.x - rho, f.ox, f.inv, f.X);
    const int x1 = (int)box_cell(qp.x + rho, f.ox, f.inv, f.X);
    const int y0 = (int)box_cell(qp.y - rho, f.oy, f.inv, f.Y);
    const int y1 = (int)box_cell(qp.y + rho, f.oy, f.inv, f.Y);
    const int z0 = (int)box_cell(qp.z - rho, f.oz, f.inv, f.Z);
    const int z1 = (int)box_cell(qp.z + rho, f.oz, f.inv, f.Z);
    const int X0 = wave_min_i(x0), X1 = wave_max_i(x1), Y0 = wave_min_i(y0),
              Y1 = wave_max_i(y1), Z0 = wave_min_i(z0), Z1 = wave_max_i(z1);
    const float cut = fminf(k1, lim);
    // the union when it is compact (its rows of cells at most twice the
    // largest box's; uniform per wave), else each lane its own box
    const int rows_own = wave_max_i((y1 - y0 + 1) * (z1 - z0 + 1));
    const bool together = (int64_t)(Y1 - Y0 + 1) * (Z1 - Z0 + 1) <= 2 * (int64_t)rows_own &&
                          X1 - X0 <= 2 * wave_max_i(x1 - x0) + 2;
    const int bx0 = together ? X0 : x0, bx1 = together ? X1 : x1;
    const int by0 = together ? Y0 : y0, by1 = together ? Y1 : y1;
    const int bz0 = together ? Z0 : z0, bz1 = together ? Z1 : z1;
    Knn<K> sel;
    sel.init();
    for (int cz = bz0; cz <= bz1; ++cz)
      for (int cy = by0; cy <= by1; ++cy) {
        const uint64_t khi = voxel_key(f, bx1, cy, cz);
        for (int32_t c = cell_lower_bound(t, voxel_key(f, bx0, cy, cz));
             c < t.ncells && ld_global(t.ukey, c) <= khi; ++c) {
          const int32_t s0 = ld_global(t.start, c), s1 = s0 + ld_global(t.count, c);
          for (int32_t j = s0; j < s1; ++j) {
            const float4 p = pts[j];
            const float pk[3] = {p.x, p.y, p.z};
            const float d = dist2(qp.x, qp.y, qp.z, pk);
            if (act && __float_as_int(p.w) != qi && d <= cut) sel.add(d, k);
          }
        }
      }
    if (act && sel.kth(k) <= lim) st_global(md, qi, sel.mean_dist(k));
  }
}

__global__ void fill_f64_kernel(double* p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) st_global(p, i, v);
}

__global__ void cloud_sor_apply_kernel(double* __restrict__ md, int64_t n,
                                       const double* __restrict__ stats,
                                       uint8_t* __restrict__ keep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = ld_global(md, i);
  if (keep) st_global(keep, i, (uint8_t)(v >= 0.0 && v <= ld_global(stats, 3) ? 1 : 0));
  if (v < 0.0) st_global(md, i, 0.0);
}

// ---------------------------------------------------------- MLS, Gauss ----

// C: the per-neighbour multiply-adds contracted (GPU semantics).
template <int NC, int TB, int R, bool C>
__global__ void __launch_bounds__(TB, 7) cloud_mls_kernel(const float4* __restrict__ pts, CellTab t,
                                                       VoxelFrame f, float r2, float inv2s2,
                                                       int order, int semantics,
                                                       float* __restrict__ out) {
  constexpr int CH = chunk_of(TB);
  __shared__ float4 tile[CH];
  __shared__ BlockCells<R> bc;
  __shared__ double etab[64];
  if (threadIdx.x < 64) etab[threadIdx.x] = kExp2Tab64[threadIdx.x];
  const XcdRange xr = xcd_range(t.ncells);
  for (int64_t c = xr.first; c < xr.end; c += xr.stride) {
    load_block<R, TB>(t, f, ld_global(t.ukey, c), &bc);  // its barriers also publish etab
    const int32_t M = bc.pre[BlockCells<R>::N];
    const int32_t c0 = ld_global(t.start, c), cn = ld_global(t.count, c);
    for (int32_t qb = 0; qb < cn; qb += TB) {
      const bool active = qb + (int32_t)threadIdx.x < cn;
      const float4 qp = active ? pts[c0 + qb + threadIdx.x] : make_float4(0.f, 0.f, 0.f, 0.f);
      const float pi[3] = {qp.x, qp.y, qp.z};
      // every lane takes part in the staging; the passes' arithmetic is the
      // organized filter's (centroid, covariance, normal equations)
      auto sweep = [&](bool on, auto&& visit) {
        for (int32_t cb = 0; cb < M; cb += CH) {
          const int32_t len = min(CH, M - cb);
          __syncthreads();
          stage_chunk<R, TB>(pts, &bc, cb, len, tile);
          __syncthreads();
          if (on)
            for (int32_t j = 0; j < len; ++j) {
              const float4 p = tile[j];
              const float pk[3] = {p.x, p.y, p.z};
              const float d2 = dist2<C>(pi[0], pi[1], pi[2], pk);
              if (d2 <= r2) visit(pk, exp_f32(-d2 * inv2s2, etab));
            }
        }
      };
      MlsFit fit;
      fit.cx = fit.cy = fit.cz = 0.f;
      float wsum = 0.f;
      int nv = 0;
      sweep(active, [&](const float pk[3], float w) {
        mls_accum_centroid<C>(pk, w, &fit, &wsum);
        nv++;
      });
      const bool go = active && nv >= 3 && !(wsum < 1e-12f);
      float o[3] = {pi[0], pi[1], pi[2]};
      float cov[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
      if (go) {
        const float iw = 1.0f / wsum;
        fit.cx *= iw;
        fit.cy *= iw;
        fit.cz *= iw;
      }
      sweep(go, [&](const float pk[3], float w) { mls_accum_cov<C>(fit, pk, w, cov); });
      int cur = 0;
      if (go) {
        mls_frame(cov, &fit);
        cur = mls_order(order, nv, semantics);
      }
      float A[NC][NC] = {}, bv[NC] = {};
      sweep(go && cur > 0, [&](const float pk[3], float w) { mls_accum_fixed<NC, C>(fit, pk, w, A, bv); });
      if (go) {
        if (cur == 0) mls_plane(fit, pi, o);
        else mls_solve_eval_fixed<NC>(fit, A, bv, cur, pi, o);
        const int qi = __float_as_int(qp.w);
        st_global(out, (int64_t)qi * 3 + 0, o[0]);
        st_global(out, (int64_t)qi * 3 + 1, o[1]);
        st_global(out, (int64_t)qi * 3 + 2, o[2]);
      }
    }
    __syncthreads();
  }
}

template <int TB, int R>
__global__ void __launch_bounds__(TB) cloud_gauss_kernel(const float4* __restrict__ pts, CellTab t,
                                                         VoxelFrame f, float r2, float inv2s2,
                                                         float* __restrict__ out) {
  constexpr int CH = chunk_of(TB);
  __shared__ float4 tile[CH];
  __shared__ BlockCells<R> bc;
  // 2^(j/256) for exp_f32_w (fewer instructions per weight than exp_f32;
  // the cloud MLS keeps exp_f32's 64-entry table, whose smaller LDS
  // footprint measures faster there)
  __shared__ double etab[256];
  for (int i = threadIdx.x; i < 256; i += TB) etab[i] = kExp2Tab256[i];
  const XcdRange xr = xcd_range(t.ncells);
  for (int64_t c = xr.first; c < xr.end; c += xr.stride) {
    load_block<R, TB>(t, f, ld_global(t.ukey, c), &bc);
    const int32_t M = bc.pre[BlockCells<R>::N];
    const int32_t c0 = ld_global(t.start, c), cn = ld_global(t.count, c);
    for (int32_t qb = 0; qb < cn; qb += TB) {
      const bool active = qb + (int32_t)threadIdx.x < cn;
      const float4 qp = active ? pts[c0 + qb + threadIdx.x] : make_float4(0.f, 0.f, 0.f, 0.f);
      float sx = 0.f, sy = 0.f, sz = 0.f, sw = 0.f;
      for (int32_t cb = 0; cb < M; cb += CH) {
        const int32_t len = min(CH, M - cb);
        __syncthreads();
        stage_chunk<R, TB>(pts, &bc, cb, len, tile);
        __syncthreads();
        if (active)
          for (int32_t j = 0; j < len; ++j) {
            const float4 p = tile[j];
            const float pk[3] = {p.x, p.y, p.z};
            const float d2 = dist2(qp.x, qp.y, qp.z, pk);
            if (d2 > r2) continue;
            const float w = exp_f32_w(-d2 * inv2s2, etab);  // as exp_f32 for d2 in [0, r2]
            sx += w * pk[0];
            sy += w * pk[1];
            sz += w * pk[2];
            sw += w;
          }
      }
      if (active) {
        const int qi = __float_as_int(qp.w);
        st_global(out, (int64_t)qi * 3 + 0, sx / sw);
        st_global(out, (int64_t)qi * 3 + 1, sy / sw);
        st_global(out, (int64_t)qi * 3 + 2, sz / sw);
      }
    }
    __syncthreads();
  }
}

// Workgroup width per cell from the mean points per occupied cell: one wave
// while cells hold tens of points (wider workgroups would leave most lanes
// idle: 2.5x slower kNN at 256 threads), wider as they fill (fewer passes of
// the block through LDS per cell).
int width_for(const CloudGridDev& g) {
  const double avg = g.ncells > 0 ? (double)g.np / g.ncells : 0.0;
  return avg > 160.0 ? 256 : (avg > 80.0 ? 128 : 64);
}

// Calls f with the workgroup width as a compile-time constant.
template <typename F>
void with_width(int tb, F&& f) {
  if (tb == 256) f(std::integral_constant<int, 256>{});
  else if (tb == 128) f(std::integral_constant<int, 128>{});
  else f(std::integral_constant<int, 64>{});
}

unsigned blocks_for(int64_t n, int per) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

// Grid over the occupied cells: up to 128 K threads' worth of workgroups per
// CU, in practice one workgroup per cell, so the hardware dispatcher balances
// cells of very different sizes (62.7 M-point merged cloud, against 2 K
// threads per CU: kNN 183 -> 147 ms, MLS 381 -> 333 ms, Gaussian 118 -> 95 ms).
int grid_for_cells(psgc_ctx* ctx, int32_t ncells, int tb) {
  const int cu = ctx->cu_count > 0 ? ctx->cu_count : 256;
  const int64_t g = std::min<int64_t>((int64_t)cu * (131072 / tb), ncells);
  return (int)std::max<int64_t>(kNumXcd, (g + kNumXcd - 1) / kNumXcd * kNumXcd);
}

}  // namespace

// The grid of the participating points of xyz (device pointers).
int cloud_grid_build(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n,
                     float voxel, CloudGridDev* g) {
  hipStream_t s = ctx->stream;
  if (n < 0 || n > ((int64_t)1 << 31) - 1) PSGC_FAIL(PSGC_E_INVALID, "cloud of %lld points", (long long)n);
  if (!(voxel > 0.0f)) PSGC_FAIL(PSGC_E_INVALID, "voxel must be > 0");
  const int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + CB - 1) / CB));
  void *misc, *k0, *k1, *i0, *i1;
  PSGC_TRY(ctx_buf(ctx, S_CL_MISC, (size_t)nb * 8 * sizeof(float) + 64, &misc));
  float* part = (float*)misc;
  if (n > 0) {
    cloud_bbox_kernel<<<nb, CB, 0, s>>>(xyz, valid, n, part);
    PSGC_LAUNCHED();
  }
  std::vector<float> hp((size_t)nb * 8, 0.f);
  if (n > 0) PSGC_HIP(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  PSGC_HIP(hipStreamSynchronize(s));
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  double np = 0;
  for (int b = 0; b < (n > 0 ? nb : 0); ++b) {
    for (int c = 0; c < 3; ++c) {
      mn[c] = std::min(mn[c], hp[b * 8 + c]);
      mx[c] = std::max(mx[c], hp[b * 8 + 3 + c]);
    }
    np += hp[b * 8 + 6];
  }
  if (np == 0) mn[0] = mn[1] = mn[2] = mx[0] = mx[1] = mx[2] = 0.f;
  if (!voxel_frame(mn, mx, voxel, &g->f))
    PSGC_FAIL(PSGC_E_INVALID, "voxel %g too small for the cloud's extent", (double)voxel);
  g->np = (int32_t)np;
  g->ncells = 0;
  if (g->np == 0) return PSGC_OK;
  const int bits = voxel_key_bits(g->f) + 1;  // one more: the sentinel sorts last
  const uint64_t sentinel = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
  PSGC_TRY(ctx_buf(ctx, S_CL_KEY0, (size_t)n * 8, &k0));
  PSGC_TRY(ctx_buf(ctx, S_CL_KEY1, (size_t)n * 8, &k1));
  PSGC_TRY(ctx_buf(ctx, S_CL_IDX0, (size_t)n * 4, &i0));
  PSGC_TRY(ctx_buf(ctx, S_CL_IDX1, (size_t)n * 4, &i1));
  cloud_keys_kernel<<<blocks_for(n, 256), 256, 0, s>>>(xyz, valid, n, g->f, sentinel, (uint64_t*)k0,
                                                       (int32_t*)i0);
  PSGC_LAUNCHED();
  // sort, run-length encode, scan; one scratch block sized for the largest
  size_t tb_sort = 0, tb_rle = 0, tb_scan = 0;
  PSGC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb_sort, (uint64_t*)k0, (uint64_t*)k1,
                                              (int32_t*)i0, (int32_t*)i1, (int)n, 0, bits, s));
  void *uk, *st, *ct;
  PSGC_TRY(ctx_buf(ctx, S_CL_UKEY, (size_t)g->np * 8, &uk));
  PSGC_TRY(ctx_buf(ctx, S_CL_START, (size_t)g->np * 4 + 4, &st));
  PSGC_TRY(ctx_buf(ctx, S_CL_COUNT, (size_t)g->np * 4 + 4, &ct));
  int32_t* nruns = (int32_t*)((char*)misc + (size_t)nb * 8 * sizeof(float));
  PSGC_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb_rle, (uint64_t*)k1, (uint64_t*)uk,
                                                 (int32_t*)ct, nruns, g->np, s));
  PSGC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb_scan, (int32_t*)ct, (int32_t*)st, g->np, s));
  void* tmp;
  size_t tb = std::max(tb_sort, std::max(tb_rle, tb_scan));
  PSGC_TRY(ctx_buf(ctx, S_CL_TEMP, tb, &tmp));
  PSGC_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, (uint64_t*)k0, (uint64_t*)k1, (int32_t*)i0,
                                              (int32_t*)i1, (int)n, 0, bits, s));
  tb = std::max(tb_sort, std::max(tb_rle, tb_scan));
  PSGC_HIP(hipcub::DeviceRunLengthEncode::Encode(tmp, tb, (uint64_t*)k1, (uint64_t*)uk,
                                                 (int32_t*)ct, nruns, g->np, s));
  int32_t hn = 0;
  PSGC_HIP(hipMemcpyAsync(&hn, nruns, 4, hipMemcpyDeviceToHost, s));
  PSGC_HIP(hipStreamSynchronize(s));
  g->ncells = hn;
  tb = std::max(tb_sort, std::max(tb_rle, tb_scan));
  PSGC_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, (int32_t*)ct, (int32_t*)st, hn, s));
  void* pts;
  PSGC_TRY(ctx_buf(ctx, S_CL_PTS, (size_t)g->np * sizeof(float4), &pts));
  cloud_gather_kernel<<<blocks_for(g->np, 256), 256, 0, s>>>(xyz, (const int32_t*)i1, g->np,
                                                             (float4*)pts);
  PSGC_LAUNCHED();
  g->pts = (const float4*)pts;
  g->ukey = (const uint64_t*)uk;
  g->start = (const int32_t*)st;
  g->count = (const int32_t*)ct;
  return PSGC_OK;
}

int cloud_sor_run(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n, int k,
                  float std_mul, float voxel, int rmax, uint8_t* keep, double* md, double* stats) {
  if (k < 1 || k > 64) PSGC_FAIL(PSGC_E_INVALID, "k=%d (1..64)", k);
  if (rmax < 1 || rmax > 16) PSGC_FAIL(PSGC_E_INVALID, "rmax=%d (1..16)", rmax);
  CloudGridDev g;
  PSGC_TRY(cloud_grid_build(ctx, xyz, valid, n, voxel, &g));
  hipStream_t s = ctx->stream;
  if (n > 0) {
    fill_f64_kernel<<<blocks_for(n, 256), 256, 0, s>>>(md, n, -1.0);
    PSGC_LAUNCHED();
  }
  if (g.ncells > 0) {
    // the retry list: count, grid positions, the block's k-th distance per
    // grid position, the flags of the first pass
    void* rl;
    PSGC_TRY(ctx_buf(ctx, S_CL_RETRY, (size_t)g.np * 9 + 64, &rl));
    int32_t* nretry = (int32_t*)rl;
    int32_t* retry = nretry + 16;
    float* kth_of = (float*)(retry + g.np);
    uint8_t* again = (uint8_t*)(kth_of + g.np);
    PSGC_HIP(hipMemsetAsync(again, 0, (size_t)g.np, s));
    const CellTab t{g.ukey, g.start, g.count, g.ncells};
    const int tbv = width_for(g);
    const int grid = grid_for_cells(ctx, g.ncells, tbv);
    const float lim = cloud_knn_limit(g.f.voxel, 1), limmax = cloud_knn_limit(g.f.voxel, rmax);
    PSGC_PROF_BEGIN("cloud_knn", s);
    with_width(tbv, [&](auto w) {
      constexpr int TB = decltype(w)::value;
      if (k == 20) cloud_knn_kernel<20, TB><<<grid, TB, 0, s>>>(g.pts, t, g.f, k, lim, limmax, md, again, kth_of);
      else if (k == 30) cloud_knn_kernel<30, TB><<<grid, TB, 0, s>>>(g.pts, t, g.f, k, lim, limmax, md, again, kth_of);
      else if (k == 50) cloud_knn_kernel<50, TB><<<grid, TB, 0, s>>>(g.pts, t, g.f, k, lim, limmax, md, again, kth_of);
      else cloud_knn_kernel<0, TB><<<grid, TB, 0, s>>>(g.pts, t, g.f, k, lim, limmax, md, again, kth_of);
    });
    PSGC_LAUNCHED();
    PSGC_PROF_END("cloud_knn", s);
    if (rmax >= 2) {
      // flagged grid positions in ascending order
      size_t tb = 0;
      const hipcub::CountingInputIterator<int32_t> ids(0);
      PSGC_HIP(hipcub::DeviceSelect::Flagged(nullptr, tb, ids, again, retry, nretry, (int)g.np, s));
      void* tmp;
      PSGC_TRY(ctx_buf(ctx, S_CL_TEMP, tb, &tmp));
      PSGC_HIP(hipcub::DeviceSelect::Flagged(tmp, tb, ids, again, retry, nretry, (int)g.np, s));
      const int rg = 4 * (ctx->cu_count > 0 ? ctx->cu_count : 256);
      PSGC_PROF_BEGIN("cloud_knn_retry", s);
      if (k == 20) cloud_knn_retry_kernel<20><<<rg, CB, 0, s>>>(g.pts, t, g.f, k, rmax, retry, kth_of, nretry, md);
      else if (k == 30) cloud_knn_retry_kernel<30><<<rg, CB, 0, s>>>(g.pts, t, g.f, k, rmax, retry, kth_of, nretry, md);
      else if (k == 50) cloud_knn_retry_kernel<50><<<rg, CB, 0, s>>>(g.pts, t, g.f, k, rmax, retry, kth_of, nretry, md);
      else cloud_knn_retry_kernel<0><<<rg, CB, 0, s>>>(g.pts, t, g.f, k, rmax, retry, kth_of, nretry, md);
      PSGC_LAUNCHED();
      PSGC_PROF_END("cloud_knn_retry", s);
    }
  }
  if (n == 0) {  // no points: zero statistics
    PSGC_HIP(hipMemsetAsync(stats, 0, 4 * sizeof(double), s));
    return PSGC_OK;
  }
  void* part;
  PSGC_TRY(ctx_buf(ctx, S_PARTIAL, sor_partial_doubles((size_t)n) * sizeof(double), &part));
  PSGC_TRY(launch_sor_stats(md, (size_t)n, std_mul, (double*)part, stats, s));
  {
    cloud_sor_apply_kernel<<<blocks_for(n, 256), 256, 0, s>>>(md, n, stats, keep);
    PSGC_LAUNCHED();
  }
  return PSGC_OK;
}

int cloud_tune_voxel(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n, int k,
                     int rmax, const float* candidates, int ncand, float* best_voxel,
                     float* times_ms) {
  if (ncand < 1 || !candidates || !best_voxel) PSGC_FAIL(PSGC_E_INVALID, "no candidates");
  hipStream_t s = ctx->stream;
  void *md, *st;
  PSGC_TRY(ctx_buf(ctx, S_CL_TUNE_MD, (size_t)std::max<int64_t>(n, 1) * 8, &md));
  PSGC_TRY(ctx_buf(ctx, S_CL_TUNE_ST, 64, &st));
  hipEvent_t a, b;
  PSGC_HIP(hipEventCreate(&a));
  PSGC_HIP(hipEventCreate(&b));
  float best = 0.f, best_ms = 1e30f;
  int rc = PSGC_OK;
  for (int c = 0; c < ncand && rc == PSGC_OK; ++c) {
    // one timed SOR per candidate (grid build included), after the first
    // candidate's untimed warm-up run
    if (c == 0) rc = cloud_sor_run(ctx, xyz, valid, n, k, 1.0f, candidates[0], rmax, nullptr,
                                   (double*)md, (double*)st);
    if (rc != PSGC_OK) break;
    if (hipEventRecord(a, s) != hipSuccess) { rc = PSGC_E_HIP; break; }
    rc = cloud_sor_run(ctx, xyz, valid, n, k, 1.0f, candidates[c], rmax, nullptr, (double*)md,
                       (double*)st);
    if (rc != PSGC_OK) break;
    float ms = 0.f;
    if (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
        hipEventElapsedTime(&ms, a, b) != hipSuccess) {
      rc = PSGC_E_HIP;
      break;
    }
    if (times_ms) times_ms[c] = ms;
    if (ms < best_ms) {
      best_ms = ms;
      best = candidates[c];
    }
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (rc != PSGC_OK) return rc;
  *best_voxel = best;
  return PSGC_OK;
}

int cloud_mls_run(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n, float sigma,
                  int order, float max_radius, int iterations, int semantics, float voxel,
                  float* out, float* tmp) {
  const float r = (max_radius > 0.0f) ? max_radius : 3.0f * sigma;
  if (!(sigma > 0.0f)) PSGC_FAIL(PSGC_E_INVALID, "sigma must be > 0");
  const int R = cloud_radius_cells(&voxel, r);
  if (R == 0) PSGC_FAIL(PSGC_E_INVALID, "voxel %g below half the radius %g", (double)voxel, (double)r);
  const float r2 = r * r, inv2s2 = 1.0f / (2.0f * sigma * sigma);
  const int ord = std::max(0, std::min(2, order));
  hipStream_t s = ctx->stream;
  const int nit = std::max(1, iterations);
  // iteration i reads src and writes dst (out last); every point starts as
  // its input, so points the fit skips pass through
  const float* src = xyz;
  for (int it = 0; it < nit; ++it) {
    float* dst = ((nit - 1 - it) % 2 == 0) ? out : tmp;
    CloudGridDev g;
    PSGC_TRY(cloud_grid_build(ctx, src, valid, n, voxel, &g));
    if (n > 0) PSGC_HIP(hipMemcpyAsync(dst, src, (size_t)n * 12, hipMemcpyDeviceToDevice, s));
    if (g.ncells > 0) {
      const CellTab t{g.ukey, g.start, g.count, g.ncells};
      // one wave per cell measures fastest for MLS at every density tried
      // (its per-candidate work dwarfs the repeated staging)
      constexpr int TB = 64;
      const int grid = grid_for_cells(ctx, g.ncells, TB);
      PSGC_PROF_BEGIN("cloud_mls", s);
      auto launch = [&](auto nc, auto rc, auto cc) {
        cloud_mls_kernel<decltype(nc)::value, TB, decltype(rc)::value, decltype(cc)::value>
            <<<grid, TB, 0, s>>>(g.pts, t, g.f, r2, inv2s2, ord, semantics, dst);
      };
      auto by_cells = [&](auto nc, auto cc) {
        if (R == 1) launch(nc, std::integral_constant<int, 1>{}, cc);
        else launch(nc, std::integral_constant<int, 2>{}, cc);
      };
      auto by_order = [&](auto cc) {
        if (ord >= 2) by_cells(std::integral_constant<int, 6>{}, cc);
        else by_cells(std::integral_constant<int, 3>{}, cc);
      };
      if (semantics == PSGC_SEM_GPU) by_order(std::true_type{});
      else by_order(std::false_type{});
      PSGC_LAUNCHED();
      PSGC_PROF_END("cloud_mls", s);
    }
    src = dst;
  }
  return PSGC_OK;
}

int cloud_gaussian_run(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n,
                       float sigma, float radius, float voxel, float* out) {
  if (!(sigma > 0.0f) || !(radius > 0.0f)) PSGC_FAIL(PSGC_E_INVALID, "sigma and radius must be > 0");
  const int R = cloud_radius_cells(&voxel, radius);
  if (R == 0)
    PSGC_FAIL(PSGC_E_INVALID, "voxel %g below half the radius %g", (double)voxel, (double)radius);
  CloudGridDev g;
  PSGC_TRY(cloud_grid_build(ctx, xyz, valid, n, voxel, &g));
  hipStream_t s = ctx->stream;
  if (n > 0) PSGC_HIP(hipMemcpyAsync(out, xyz, (size_t)n * 12, hipMemcpyDeviceToDevice, s));
  if (g.ncells > 0) {
    const CellTab t{g.ukey, g.start, g.count, g.ncells};
    PSGC_PROF_BEGIN("cloud_gauss", s);
    // at least two waves on radius cells (their blocks are long to stage);
    // on half-radius cells, which hold a quarter of the points of a surface,
    // the density rule alone (62.7 M-point merged cloud: 129 -> 117 ms)
    const int tbv = R == 2 ? width_for(g) : std::max(128, width_for(g));
    with_width(tbv, [&](auto w) {
      constexpr int TB = decltype(w)::value;
      const unsigned grid = grid_for_cells(ctx, g.ncells, TB);
      const float r2 = radius * radius, inv2s2 = 1.0f / (2.0f * sigma * sigma);
      if (R == 1) cloud_gauss_kernel<TB, 1><<<grid, TB, 0, s>>>(g.pts, t, g.f, r2, inv2s2, out);
      else cloud_gauss_kernel<TB, 2><<<grid, TB, 0, s>>>(g.pts, t, g.f, r2, inv2s2, out);
    });
    PSGC_LAUNCHED();
    PSGC_PROF_END("cloud_gauss", s);
  }
  return PSGC_OK;
}

}  // namespace psgc
