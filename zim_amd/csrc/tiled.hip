// tiled.hip — LDS-tiled, batch-capable kernels of the pipeline.
//
// Same per-pixel arithmetic as the plain kernels (psgc_device.h, reference
// CPU operation order, no contraction) so the integer outputs stay
// bit-identical; what changes is the data movement:
//   decode     one thread per column, 4 rows per thread, each raw frame tile
//              (6 x 258 samples) staged once through LDS (double-buffered),
//              the horizontal blur pass done once per sample;
//   match      one workgroup per row: the right phase row (post-smoothed on
//              the fly from the raw map) and its 16-bit orders in LDS, per-row order
//              spans built with LDS atomics, the argmin scan reads LDS only;
//   median     20x20 tile in LDS, branch-free 202-op selection network;
//   reproject+Gaussian fused on a 16x16 tile with its halo in LDS;
//   SOR        k-nearest search over the LDS tile, ring order (near first);
//   MLS        points and mask of the tile plus halo in LDS, the in-radius
//              window entries kept as a bit mask so passes 2 and 3 skip the
//              distance test, fast exp2-based weight.

#include <climits>

#include "psgc_device.h"
#include "psgc_internal.h"
#include "psgc_math.h"
#include "psgc_tiled.h"
#include "psgc_select.h"

namespace psgc {
namespace {

// ----------------------------------------------------------------- decode --

constexpr int DC = 256;  // tile width (threads)
constexpr int DLC = DC + 2;

// A tile is DC columns by DRT rows (DRT output rows per thread, chosen by the
// batch autotuner); its frame samples span (DRT + 2) x (DC + 2) with the
// replicated border. Thread t holds column t of every tile row; threads
// t < 2 (DRT + 2) also hold one of the samples past column DC (row t / 2,
// column DC + t % 2).
template <int DRT>
struct FrameRegs {
  double v[DRT + 2];
  double e;
};

template <int DRT>
constexpr int decode_extra() { return (DLC - DC) * (DRT + 2); }

// f64 frames: decode_fetch issues the next frame's loads into registers,
// decode_store writes them to LDS after the current frame is computed.
template <int DRT>
__device__ inline void decode_fetch(const DecodeArgs& a, int f, int y0, int x0,
                                    FrameRegs<DRT>* r) {
  const int t = threadIdx.x;
  const int gx = min(max(x0 - 1 + t, 0), a.cols - 1);
#pragma unroll
  for (int rr = 0; rr < DRT + 2; ++rr) {
    const int gy = min(max(y0 - 1 + rr, 0), a.rows - 1);
    r->v[rr] = load_sample(a, f, gy, gx);
  }
  if (t < decode_extra<DRT>()) {
    const int gy = min(max(y0 - 1 + t / 2, 0), a.rows - 1);
    const int gxe = min(max(x0 - 1 + DC + t % 2, 0), a.cols - 1);
    r->e = load_sample(a, f, gy, gxe);
  }
}

template <int DRT>
__device__ inline void decode_store(const FrameRegs<DRT>& r, double (*buf)[DLC]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int rr = 0; rr < DRT + 2; ++rr) buf[rr][t] = r.v[rr];
  if (t < decode_extra<DRT>()) buf[t / 2][DC + t % 2] = r.e;
}

// Integer frames (u8, u16) are staged DCH frames at a time as raw 16-bit
// samples, with one barrier per chunk instead of one per frame; f64 frames
// use two double buffers. Both layouts share one LDS block.
constexpr int DCH = 12;

template <int DRT>
constexpr size_t decode_lds_bytes() {
  const size_t raw = (size_t)DCH * (DRT + 2) * DLC * sizeof(uint16_t);
  const size_t f64 = (size_t)2 * (DRT + 2) * DLC * sizeof(double);
  return raw > f64 ? raw : f64;
}

__device__ inline uint16_t load_raw16(const DecodeArgs& a, int f, int r, int x) {
  const char* row = (const char*)a.frames[f] + (size_t)r * a.stride;
  if (a.sample_type == PSGC_SAMPLE_U16) return ld_global((const uint16_t*)row, x);
  return ld_global((const uint8_t*)row, x);
}

// Frame f's (DRT + 2) x (DC + 2) tile (replicated border) as raw samples.
template <int DRT>
__device__ inline void decode_stage_raw(const DecodeArgs& a, int f, int y0, int x0,
                                        uint16_t (*raw)[DLC]) {
  const int t = threadIdx.x;
  const int gx = min(max(x0 - 1 + t, 0), a.cols - 1);
  uint16_t v[DRT + 2];
#pragma unroll
  for (int rr = 0; rr < DRT + 2; ++rr)
    v[rr] = load_raw16(a, f, min(max(y0 - 1 + rr, 0), a.rows - 1), gx);
  uint16_t e = 0;
  if (t < decode_extra<DRT>())
    e = load_raw16(a, f, min(max(y0 - 1 + t / 2, 0), a.rows - 1),
                   min(max(x0 - 1 + DC + t % 2, 0), a.cols - 1));
#pragma unroll
  for (int rr = 0; rr < DRT + 2; ++rr) raw[rr][t] = v[rr];
  if (t < decode_extra<DRT>()) raw[t / 2][DC + t % 2] = e;
}

// Frames f0 .. f0 + nf - 1 (nf <= DCH) of the tile as raw samples of type S,
// the loads of DSB frames issued before their LDS writes, so that a tile
// waits on DCH / DSB round trips to memory rather than one per frame. Used
// for 2-row tiles (measured 5 % faster than frame by frame; 4-row tiles hold
// the level sums of twice the pixels, and more loads in flight spill
// registers there).
constexpr int DSB = 6;
static_assert(DCH % DSB == 0, "whole sub-batches");

template <int DRT, typename S>
__device__ inline void decode_stage_chunk(const DecodeArgs& a, int f0, int nf, int y0, int x0,
                                          uint16_t (*raw)[DRT + 2][DLC]) {
  const int t = threadIdx.x;
  const int gx = min(max(x0 - 1 + t, 0), a.cols - 1);
  const bool ext = t < decode_extra<DRT>();
  const int gxe = min(max(x0 - 1 + DC + t % 2, 0), a.cols - 1);
  const size_t rowe = (size_t)min(max(y0 - 1 + t / 2, 0), a.rows - 1) * a.stride;
  size_t row[DRT + 2];
#pragma unroll
  for (int rr = 0; rr < DRT + 2; ++rr)
    row[rr] = (size_t)min(max(y0 - 1 + rr, 0), a.rows - 1) * a.stride;
  for (int g0 = 0; g0 < nf; g0 += DSB) {
    uint16_t v[DSB][DRT + 2], e[DSB];
#pragma unroll
    for (int g = 0; g < DSB; ++g) {
      if (g0 + g < nf) {  // uniform
        const char* fr = (const char*)a.frames[f0 + g0 + g];
#pragma unroll
        for (int rr = 0; rr < DRT + 2; ++rr) v[g][rr] = ld_global((const S*)(fr + row[rr]), gx);
        e[g] = ext ? ld_global((const S*)(fr + rowe), gxe) : (uint16_t)0;
      }
    }
#pragma unroll
    for (int g = 0; g < DSB; ++g) {
      if (g0 + g < nf) {
#pragma unroll
        for (int rr = 0; rr < DRT + 2; ++rr) raw[g0 + g][rr][t] = v[g][rr];
        if (ext) raw[g0 + g][t / 2][DC + t % 2] = e[g];
      }
    }
  }
}

// 4-row tiles: four waves per SIMD (at most 128 VGPRs) measure faster than
// three without the bound, despite a few spilled scalars. 2-row tiles: five
// (at most 96 VGPRs, 36 bytes per lane spilled) measure 2-3 % faster than
// four; six spill more and measure slower.
template <int DRT>
__global__ void __launch_bounds__(DC, DRT <= 2 ? 5 : 4)
    decode_tiled_kernel(const DecodeJob* __restrict__ jobs, int njobs, int64_t total) {
  constexpr int DLR = DRT + 2;
  __shared__ __align__(16) unsigned char lds[decode_lds_bytes<DRT>()];
  // the atan2 table (published by the first tile's staging barrier)
  __shared__ double atab[65];
  if (threadIdx.x < 65) atab[threadIdx.x] = kAtanTab[threadIdx.x];
  const XcdRange xr = xcd_range(total);
  int jprev = 0;
  for (int64_t t = xr.first; t < xr.end; t += xr.stride) {
    const int j = find_job(jobs, njobs, t, jprev);
    jprev = j;
    const DecodeJob& job = jobs[j];
    const DecodeArgs& a = job.a;
    const int64_t lt = t - job.tile_base;
    const int tx = (int)(lt % job.tiles_x), ty = (int)(lt / job.tiles_x);
    const int x0 = tx * DC, y0 = ty * DRT;
    const int x = x0 + threadIdx.x;
    const double c = a.blur_c, s = a.blur_s;
    const int tc = threadIdx.x + 1;

    double S[DRT], Cs[DRT], G[DRT], wrap[3][DRT], mod0[DRT];
    bool ok[DRT];
    for (int q = 0; q < DRT; ++q) {
      ok[q] = true;
      mod0[q] = 0.0;
      wrap[0][q] = wrap[1][q] = wrap[2][q] = 0.0;
      S[q] = Cs[q] = G[q] = 0.0;
    }
    int nframes = 0;
    for (int i = 0; i < a.nlev; ++i) nframes += a.shift_num[i];
    int level = 0, kin = 0;

    // one frame's blurred samples into the level sums (h: the horizontal
    // pass of the tile rows at this thread's column)
    auto accumulate = [&](int f, const double (&h)[DLR]) {
      const double st = a.sin_tab[f], ct = a.cos_tab[f];
#pragma unroll
      for (int q = 0; q < DRT; ++q) {
        const double v = c * h[q + 1] + s * (h[q] + h[q + 2]);
        S[q] -= st * v;
        Cs[q] += ct * v;
        G[q] += v;
      }
      if (++kin == a.shift_num[level]) {
#pragma unroll
        for (int q = 0; q < DRT; ++q) {
          double m;
          const bool lv = level_finish(S[q], Cs[q], G[q], a.shift_num[level], a.mod_thresh,
                                       &wrap[level][q], &m, atab);
          if (level == 0) mod0[q] = m;
          ok[q] = ok[q] && lv;
          S[q] = Cs[q] = G[q] = 0.0;
        }
        kin = 0;
        ++level;
      }
    };

    if (a.sample_type != PSGC_SAMPLE_F64) {
      uint16_t(*raw)[DLR][DLC] = (uint16_t(*)[DLR][DLC])lds;
      for (int f0 = 0; f0 < nframes; f0 += DCH) {
        const int nf = min(DCH, nframes - f0);
        if (DRT > 2)
          for (int g = 0; g < nf; ++g) decode_stage_raw<DRT>(a, f0 + g, y0, x0, raw[g]);
        else if (a.sample_type == PSGC_SAMPLE_U16)
          decode_stage_chunk<DRT, uint16_t>(a, f0, nf, y0, x0, raw);
        else
          decode_stage_chunk<DRT, uint8_t>(a, f0, nf, y0, x0, raw);
        __syncthreads();
        for (int g = 0; g < nf; ++g) {
          const uint16_t(*b)[DLC] = raw[g];
          double h[DLR];
#pragma unroll
          for (int rr = 0; rr < DLR; ++rr)
            // integer samples: the outer taps' sum is exact in integers
            // (and in double), so it is converted once
            h[rr] = c * (double)b[rr][tc] + s * (double)((int)b[rr][tc - 1] + (int)b[rr][tc + 1]);
          accumulate(f0 + g, h);
        }
        __syncthreads();  // the chunk's readers are done before it is refilled
      }
    } else {
      double(*buf)[DLR][DLC] = (double(*)[DLR][DLC])lds;
      FrameRegs<DRT> pre;
      decode_fetch<DRT>(a, 0, y0, x0, &pre);
      decode_store<DRT>(pre, buf[0]);
      __syncthreads();
      for (int f = 0; f < nframes; ++f) {
        if (f + 1 < nframes) decode_fetch<DRT>(a, f + 1, y0, x0, &pre);
        const double(*b)[DLC] = buf[f & 1];
        double h[DLR];
#pragma unroll
        for (int rr = 0; rr < DLR; ++rr)
          h[rr] = c * b[rr][tc] + s * (b[rr][tc - 1] + b[rr][tc + 1]);
        accumulate(f, h);
        // buf[(f + 1) & 1] was last read in iteration f - 1, before that
        // iteration's barrier
        if (f + 1 < nframes) decode_store<DRT>(pre, buf[(f + 1) & 1]);
        __syncthreads();
      }
    }
    if (x < a.cols) {
      for (int q = 0; q < DRT; ++q) {
        const int y = y0 + q;
        if (y >= a.rows) break;
        const size_t p = (size_t)y * a.cols + x;
        double ph = 0.0;
        int32_t k = 0;
        bool v = ok[q];
        if (v) {
          double w3[3] = {wrap[0][q], wrap[1][q], wrap[2][q]};
          if (a.nlev == 1) v = gray_order(a, y, x, w3[0], &ph, &k);
          else heterodyne(a, w3, &ph, &k);
        }
        if (!v) {
          ph = 0.0;
          k = 0;
        }
        st_global(job.phase, p, ph);
        st_global(job.k, p, k);
        if (job.mod) st_global(job.mod, p, mod0[q]);
        if (job.valid) st_global(job.valid, p, (uint8_t)(v ? 255 : 0));
      }
    }
  }
}

// ------------------------------------------------------------------ match --

__device__ inline double smooth_at(const double* m, int rows, int cols, int r, int x) {
  const int xl = x > 0 ? x - 1 : 0, xr = x < cols - 1 ? x + 1 : cols - 1;
  const int ru = r > 0 ? r - 1 : 0, rd = r < rows - 1 ? r + 1 : rows - 1;
  const double* u = m + (size_t)ru * cols;
  const double* c = m + (size_t)r * cols;
  const double* d = m + (size_t)rd * cols;
  const double hu = 0.5 * ld_global(u, x) + 0.25 * (ld_global(u, xl) + ld_global(u, xr));
  const double hm = 0.5 * ld_global(c, x) + 0.25 * (ld_global(c, xl) + ld_global(c, xr));
  const double hd = 0.5 * ld_global(d, x) + 0.25 * (ld_global(d, xl) + ld_global(d, xr));
  return 0.5 * hm + 0.25 * (hu + hd);
}

// One rectified row per workgroup of MT threads (the row tile the batch
// autotuner picks: 128, 256 or 512), the row's right phases and orders and
// its order spans in LDS.
template <int MT>
__global__ void __launch_bounds__(MT) match_tiled_kernel(const MatchJob* __restrict__ jobs,
                                                         int njobs, int64_t total) {
  extern __shared__ double lds[];
  const XcdRange xr = xcd_range(total);
  int jprev = 0;
  for (int64_t t = xr.first; t < xr.end; t += xr.stride) {
    const int j = find_job(jobs, njobs, t, jprev);
    jprev = j;
    const MatchJob& m = jobs[j];
    const int r = (int)(t - m.tile_base);
    const int cols = m.cols;
    const int knum = m.pha_f > 0 ? m.pha_f : 0;
    // the right row's orders as 16 bits (orders run to pha_f <= 32767):
    // 10 bytes per pixel, so three workgroups of a 5120-px row share a CU
    double* sR = lds;
    int16_t* kR = (int16_t*)(sR + cols);
    int32_t* lo = (int32_t*)(kR + ((cols + 1) & ~1));
    int32_t* hi = lo + (knum + 1);
    const size_t row = (size_t)r * cols;
    for (int k = threadIdx.x; k <= knum; k += MT) {
      lo[k] = INT_MAX;
      hi[k] = -1;
    }
    for (int c = threadIdx.x; c < cols; c += MT) {
      sR[c] = m.smooth ? smooth_at(m.rp, m.rows, cols, r, c) : ld_global(m.rp, row + c);
      // an order outside 16 bits is never matched (above pha_f): stored as -1
      const int32_t kc = ld_global(m.rk, row + c);
      kR[c] = (int16_t)(kc >= -1 && kc <= 32767 ? kc : -1);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < cols; c += MT) {
      const int k = kR[c];
      if (k >= 0 && k <= knum) {
        atomicMin(&lo[k], c);
        atomicMax(&hi[k], c);
      }
    }
    __syncthreads();
    for (int jj = threadIdx.x; jj < cols; jj += MT) {
      double d = 0.0;
      int best = -1;
      const int k = ld_global(m.lk, row + jj);
      if (k > 0 && k <= m.pha_f) {
        const int a = lo[k], b = hi[k];
        const double lv =
            m.smooth ? smooth_at(m.lp, m.rows, cols, r, jj) : ld_global(m.lp, row + jj);
        double min_cost = 1e9;
        for (int c = a; c <= b; ++c) {
          if (kR[c] != k) continue;
          const double diff = fabs(sR[c] - lv);
          if (diff < min_cost) {
            min_cost = diff;
            best = c;
          }
        }
        if (best >= 0 && min_cost < m.grad_thr && best >= 2 && best < cols - 1) {
          d = subpixel_disparity(sR, best, lv, jj);
        } else {
          best = -1;
        }
      }
      st_global(m.disp, row + jj, (float)d);
      if (m.disp64) st_global(m.disp64, row + jj, d);
      if (m.match_col) st_global(m.match_col, row + jj, (int32_t)best);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ cloud tiles --

constexpr int CT = 16;  // 16x16 output tile, 256 threads

__device__ inline float median25(const float* v) {
  const float t0 = v[0], t1 = v[1], t2 = v[2], t3 = v[3], t4 = v[4], t5 = v[5], t6 = v[6],
              t7 = v[7], t8 = v[8], t9 = v[9], t10 = v[10], t11 = v[11], t12 = v[12],
              t13 = v[13], t14 = v[14], t15 = v[15], t16 = v[16], t17 = v[17], t18 = v[18],
              t19 = v[19], t20 = v[20], t21 = v[21], t22 = v[22], t23 = v[23], t24 = v[24];
#include "median25.inc"
  return PSGC_MEDIAN25_RESULT;
}

__global__ void __launch_bounds__(256) median_tiled_kernel(const CloudJob* __restrict__ jobs,
                                                           int njobs, int64_t total) {
  __shared__ float tile[CT + 4][CT + 5];
  const XcdRange xr = xcd_range(total);
  const int lx = threadIdx.x % CT, ly = threadIdx.x / CT;
  int jprev = 0;
  for (int64_t t = xr.first; t < xr.end; t += xr.stride) {
    const int j = find_job(jobs, njobs, t, jprev);
    jprev = j;
    const CloudJob& cj = jobs[j];
    const int64_t lt = t - cj.tile_base;
    const int x0 = (int)(lt % cj.tiles_x) * CT, y0 = (int)(lt / cj.tiles_x) * CT;
    for (int idx = threadIdx.x; idx < (CT + 4) * (CT + 4); idx += 256) {
      const int rr = idx / (CT + 4), cc = idx % (CT + 4);
      const int gy = min(max(y0 - 2 + rr, 0), cj.rows - 1);
      const int gx = min(max(x0 - 2 + cc, 0), cj.cols - 1);
      tile[rr][cc] = ld_global(cj.disp_in, (size_t)gy * cj.cols + gx);
    }
    __syncthreads();
    const int x = x0 + lx, y = y0 + ly;
    if (x < cj.cols && y < cj.rows) {
      float v[25];
#pragma unroll
      for (int dy = 0; dy < 5; ++dy)
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) v[dy * 5 + dx] = tile[ly + dy][lx + dx];
      st_global(cj.disp_med, (size_t)y * cj.cols + x, median25(v));
    }
    __syncthreads();
  }
}

constexpr int GH_MAX = 3;  // Gaussian half window <= 3 (win <= 7)
constexpr int GL = CT + 2 * GH_MAX;

__global__ void __launch_bounds__(256) repro_gauss_tiled_kernel(const CloudJob* __restrict__ jobs,
                                                                int njobs, int64_t total) {
  __shared__ float px[GL][GL + 1], py[GL][GL + 1], pz[GL][GL + 1];
  __shared__ uint8_t pm[GL][GL + 1];
  const XcdRange xr = xcd_range(total);
  const int lx = threadIdx.x % CT, ly = threadIdx.x / CT;
  int jprev = 0;
  for (int64_t t = xr.first; t < xr.end; t += xr.stride) {
    const int j = find_job(jobs, njobs, t, jprev);
    jprev = j;
    const CloudJob& cj = jobs[j];
    const int64_t lt = t - cj.tile_base;
    const int x0 = (int)(lt % cj.tiles_x) * CT, y0 = (int)(lt / cj.tiles_x) * CT;
    const int h = (cj.gwin - 1) / 2;
    const int L = CT + 2 * h;
    // the tile's disparities, both of a thread's loads issued before either
    // is used ((CT + 2 GH_MAX)^2 <= 2 x 256 entries)
    static_assert(GL * GL <= 2 * 256, "two entries per thread");
    double dds[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int idx = threadIdx.x + u * 256;
      const int rr = idx / L, cc = idx % L;
      const int gy = y0 - h + rr, gx = x0 - h + cc;
      dds[u] = 0.0;
      if (idx < L * L && gy >= 0 && gy < cj.rows && gx >= 0 && gx < cj.cols) {
        const size_t p = (size_t)gy * cj.cols + gx;
        dds[u] = cj.disp_src64 ? ld_global(cj.disp_src64, p) : (double)ld_global(cj.disp_src, p);
      }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int idx = threadIdx.x + u * 256;
      if (idx >= L * L) break;
      const int rr = idx / L, cc = idx % L;
      const int gy = y0 - h + rr, gx = x0 - h + cc;
      float o[3] = {0.f, 0.f, 0.f};
      uint8_t v = 0;
      if (gy >= 0 && gy < cj.rows && gx >= 0 && gx < cj.cols) {
        const size_t p = (size_t)gy * cj.cols + gx;
        const double dd = dds[u];
        v = cj.coef[2] ? reproject_poly(ld_global(cj.coef[0], p), ld_global(cj.coef[1], p),
                                        ld_global(cj.coef[2], p), cj.r1, dd,
                                        cj.min_depth, cj.max_depth, cj.use_rect, o)
                       : reproject_point(cj.q, cj.r1, gy, gx, dd, cj.min_depth,
                                         cj.max_depth, cj.use_rect, o);
        // the tile's own pixels publish the reprojection
        if (rr >= h && rr < h + CT && cc >= h && cc < h + CT) {
          if (cj.xyz0) {
            st_global(cj.xyz0, p * 3 + 0, o[0]);
            st_global(cj.xyz0, p * 3 + 1, o[1]);
            st_global(cj.xyz0, p * 3 + 2, o[2]);
          }
          if (cj.P0) st_global(cj.P0, p, v);
        }
      }
      px[rr][cc] = o[0];
      py[rr][cc] = o[1];
      pz[rr][cc] = o[2];
      pm[rr][cc] = v;
    }
    __syncthreads();
    const int x = x0 + lx, y = y0 + ly;
    if (x < cj.cols && y < cj.rows) {
      const size_t p = (size_t)y * cj.cols + x;
      const int cy = ly + h, cx = lx + h;
      float o0 = 0.f, o1 = 0.f, o2 = 0.f;
      uint8_t op = cj.semantics == PSGC_SEM_CPU ? pm[cy][cx] : 0;
      if (y >= h && y < cj.rows - h && x >= h && x < cj.cols - h && pm[cy][cx] != 0) {
        const float ccx = px[cy][cx];
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, cnt = 0.f;
        for (int ik = -h; ik <= h; ++ik)
          for (int jk = -h; jk <= h; ++jk) {
            const int ry = cy + ik, rx = cx + jk;
            if (pm[ry][rx] == 0) continue;
            const float qx = px[ry][rx];
            if (fabsf(qx - ccx) < 3.0f) {
              const float w = cj.gk[(ik + h) * cj.gwin + (jk + h)];
              s0 += qx * w;
              s1 += py[ry][rx] * w;
              s2 += pz[ry][rx] * w;
              cnt += w;
            }
          }
        if (cnt > 0.8f) {
          if (cj.semantics == PSGC_SEM_CPU) {
            const float inv = 1.f / cnt;
            o0 = s0 * inv;
            o1 = s1 * inv;
            o2 = s2 * inv;
          } else {
            o0 = s0 / cnt;
            o1 = s1 / cnt;
            o2 = s2 / cnt;
            op = 1;
          }
        }
      }
      st_global(cj.xyzg, p * 3 + 0, o0);
      st_global(cj.xyzg, p * 3 + 1, o1);
      st_global(cj.xyzg, p * 3 + 2, o2);
      st_global(cj.Pg, p, op);
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------- SOR --

constexpr int SH_MAX = kSorTiledMaxHalf;
constexpr int SL = CT + 2 * SH_MAX;

// Loads points and mask of the tile plus a halo of `h` into LDS (mask 0
// outside the image).
template <int LD>
__device__ inline void load_cloud_tile(const float* xyz, const uint8_t* P, const double* md,
                                       const double* thr, int rows, int cols, int y0, int x0,
                                       int h, float4 (*T)[LD + 1]) {
  // T[r][c] = (x, y, z, 1) for a valid point, (x, y, z, 0) otherwise, so one
  // 16-byte LDS read returns a point and its mask
  const int L = CT + 2 * h;
  const double th = thr ? ld_global(thr, 0) : 0.0;
  for (int idx = threadIdx.x; idx < L * L; idx += 256) {
    const int rr = idx / L, cc = idx % L;
    const int gy = y0 - h + rr, gx = x0 - h + cc;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (gy >= 0 && gy < rows && gx >= 0 && gx < cols) {
      const size_t p = (size_t)gy * cols + gx;
      bool ok = ld_global(P, p) != 0;
      if (ok && md) {
        const double d = ld_global(md, p);
        ok = d >= 0.0 && d <= th;
      }
      v = make_float4(ld_global(xyz, p * 3 + 0), ld_global(xyz, p * 3 + 1),
                      ld_global(xyz, p * 3 + 2), ok ? 1.f : 0.f);
    }
    T[rr][cc] = v;
  }
}

// Keeps the KMAX smallest values seen (ascending in best[0..cnt)), capacity k.
template <int KMAX>
__device__ inline void keep_smallest(float (&best)[KMAX], int& cnt, int k, float v) {
  if (cnt == k && !(v < best[k - 1])) return;
#pragma unroll
  for (int t = 0; t < KMAX; ++t) {
    if (t < cnt) {
      const float b = best[t];
      const bool sw = v < b;
      best[t] = sw ? v : b;
      v = sw ? b : v;
    } else if (t == cnt) {
      best[t] = v;
    }
  }
  if (cnt < k) ++cnt;
}

// Exactly K smallest, list prefilled with +inf: every insertion is the same
// branch-free carry chain; the point is eligible iff the list filled up.
// Keeps the K smallest values in ascending order: a new value replaces the
// largest and bubbles down until it is in place. The window is scanned in
// rings of growing radius, so most values land at or near the end and the
// bubble stops after a step or two (a wave leaves the chain when all its
// lanes have).
template <int K>
__device__ inline void keep_smallest_fixed(float (&best)[K], float v) {
  if (!(v < best[K - 1])) return;
  best[K - 1] = v;
#pragma unroll
  for (int t = K - 1; t > 0; --t) {
    const float a = best[t - 1], b = best[t];
    if (!(b < a)) break;
    best[t - 1] = b;
    best[t] = a;
  }
}

// Visits the window entries at Chebyshev distance 1..h from (cy, cx), ring by
// ring (near first), each ring as four runs: the top row left to right, the
// right column downwards, the bottom row right to left, the left column
// upwards. Loop bounds are uniform, so no per-entry index arithmetic.
template <typename F>
__device__ inline void for_each_ring_entry(int h, int cy, int cx, F&& visit, int rad0 = 1) {
  for (int rad = rad0; rad <= h; ++rad) {
    for (int o = -rad; o < rad; ++o) visit(cy - rad, cx + o);
    for (int o = -rad; o < rad; ++o) visit(cy + o, cx + rad);
    for (int o = rad; o > -rad; --o) visit(cy + rad, cx + o);
    for (int o = rad; o > -rad; --o) visit(cy + o, cx - rad);
  }
}

// KFIX > 0: k == KFIX known at compile time; otherwise k <= KMAX at run time.
template <int KMAX, int KFIX>
__global__ void __launch_bounds__(256) sor_knn_tiled_kernel(const CloudJob* __restrict__ jobs,
                                                            int njobs, int64_t total) {
  __shared__ float4 T[SL][SL + 1];
  const XcdRange xr = xcd_range(total);
  const int lx = threadIdx.x % CT, ly = threadIdx.x / CT;
  int jprev = 0;
  for (int64_t t = xr.first; t < xr.end; t += xr.stride) {
    const int j = find_job(jobs, njobs, t, jprev);
    jprev = j;
    const CloudJob& cj = jobs[j];
    const int64_t lt = t - cj.tile_base;
    const int x0 = (int)(lt % cj.tiles_x) * CT, y0 = (int)(lt / cj.tiles_x) * CT;
    const int h = cj.sor_half;
    load_cloud_tile<SL>(cj.sor_xyz, cj.sor_P, nullptr, nullptr, cj.rows, cj.cols, y0, x0, h, T);
    __syncthreads();
    const int x = x0 + lx, y = y0 + ly;
    if (x < cj.cols && y < cj.rows) {
      const int cy = ly + h, cx = lx + h;
      double out = -1.0;
      const float4 q = T[cy][cx];
      if (q.w != 0.f) {
        const float qx = q.x, qy = q.y, qz = q.z;
        if constexpr (KFIX > 0) {
          // Rings 1 and 2 (24 entries; invalid ones +inf) sorted by a fixed
          // network, their KFIX smallest kept; the rest of the window enters
          // one by one. The selected multiset, and so the ascending list, is
          // the same in any order. A window of fewer than KFIX entries never
          // holds KFIX neighbours (and KFIX > 8 means such a window has h < 2).
          static_assert(KFIX > 8, "the fixed-k selection prefills rings 1-2");
          constexpr int NP = 24;
          constexpr int NB = KFIX > NP ? KFIX : NP;
          const bool enough = (2 * h + 1) * (2 * h + 1) - 1 >= KFIX;
          float cand[NB];
#pragma unroll
          for (int i = 0; i < NB; ++i) cand[i] = __builtin_huge_valf();
          if (enough) {
#pragma unroll
            for (int i = 0; i < NP; ++i) {
              // ring(i) <= 2 <= h: inside the tile's halo
              const float4 e4 = T[cy + kRing.dy[i]][cx + kRing.dx[i]];
              const float p3[3] = {e4.x, e4.y, e4.z};
              const float d2 = dist2(qx, qy, qz, p3);
              // a NaN distance never enters the list (as with the insertion below)
              cand[i] = ((e4.w != 0.f) & (d2 < __builtin_huge_valf())) ? d2 : __builtin_huge_valf();
            }
          }
          constexpr SortNet net = make_sort_net(NP);
#pragma unroll
          for (int c = 0; c < net.n; ++c) {
            const float a = cand[net.a[c]], b = cand[net.b[c]];
            cand[net.a[c]] = __builtin_fminf(a, b);
            cand[net.b[c]] = __builtin_fmaxf(a, b);
          }
          float best[KFIX];
#pragma unroll
          for (int i = 0; i < KFIX; ++i) best[i] = cand[i];
          for_each_ring_entry(
              enough ? h : 0, cy, cx,
              [&](int ry, int rx) {
                const float4 e4 = T[ry][rx];
                if (e4.w == 0.f) return;
                const float p3[3] = {e4.x, e4.y, e4.z};
                keep_smallest_fixed<KFIX>(best, dist2(qx, qy, qz, p3));
              },
              3);
          if (best[KFIX - 1] < __builtin_huge_valf()) {
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < KFIX; ++m) s += sqrt_of_float(best[m]);
            out = s / KFIX;
          }
        } else {
          float best[KMAX];
          int cnt = 0;
          const int k = cj.sor_k;
          for_each_ring_entry(h, cy, cx, [&](int ry, int rx) {
            const float4 e4 = T[ry][rx];
            if (e4.w == 0.f) return;
            const float p3[3] = {e4.x, e4.y, e4.z};
            keep_smallest<KMAX>(best, cnt, k, dist2(qx, qy, qz, p3));
          });
          if (cnt >= k) {
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < KMAX; ++m)
              if (m < k) s += sqrt_of_float(best[m]);
            out = s / k;
          }
        }
      }
      st_global(cj.md, (size_t)y * cj.cols + x, out);
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------- MLS --

static_assert(kMlsClassHalf[kMlsClasses - 1] == kMlsTiledMaxHalf, "classes cover the limit");
// Weight-cache entries per thread of each window class: as many as keep four
// workgroups per CU resident beside the class's LDS tile (tile + cache about
// 37 KB), the occupancy the registers allow; none above half window 10.
template <int HM>
constexpr int mls_cache_entries() {
  return HM <= 6 ? 24 : (HM <= 8 ? 20 : (HM <= 10 ? 16 : 0));
}

// Visits the in-radius entries of window rows [dy0, dy1) (bit masks per row)
// in order, with their weights: from the cache (entries kc, kc + 1, ... of
// this thread's column) when CACHED, else the pass-1 expression evaluated
// again (bit-identical either way).
template <bool CACHED, int LD, bool C, typename F>
__device__ inline void mls_rows(const uint32_t* bits, int dy0, int dy1, int ry0, int rx0,
                                const float4 (*T)[LD + 1], const float (*Wc)[256], int& kc,
                                const double* etab, const float pi[3], float inv2s2, F&& visit) {
  for (int dy = dy0; dy < dy1; ++dy) {
    const int ry = ry0 + dy;
    uint32_t b = bits[dy];
    while (b) {
      const float4 e4 = T[ry][rx0 + __builtin_ctz(b)];
      b &= b - 1;
      const float pk[3] = {e4.x, e4.y, e4.z};
      float w;
      if (CACHED) w = Wc[kc][threadIdx.x];
      else w = exp_f32_w(-dist2<C>(pi[0], pi[1], pi[2], pk) * inv2s2, etab);
      ++kc;
      visit(pk, w);
    }
  }
}

// All in-radius entries of the window in order: rows [0, dyc) from the cache.
template <int LD, bool C, typename F>
__device__ inline void mls_entries(const uint32_t* bits, int dyc, int Wd, int ry0, int rx0,
                                   const float4 (*T)[LD + 1], const float (*Wc)[256],
                                   const double* etab, const float pi[3], float inv2s2,
                                   F&& visit) {
  int kc = 0;
  if (dyc > 0) mls_rows<true, LD, C>(bits, 0, dyc, ry0, rx0, T, Wc, kc, etab, pi, inv2s2, visit);
  mls_rows<false, LD, C>(bits, dyc, Wd, ry0, rx0, T, Wc, kc, etab, pi, inv2s2, visit);
}

__device__ inline int wave_min(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = min(v, __shfl_xor(v, m, 64));
  return v;
}

// HM: largest half window served; KC: weight-cache entries per thread (0:
// every pass evaluates its weights); C: the per-neighbour multiply-adds
// contracted (GPU semantics); tiles of jobs of the other semantics are left
// to the launch for those. The classes above half window 10 are held to four
// waves per SIMD (at most 128 VGPRs, 20-28 bytes per lane spilled), which
// measured 9 % faster on an uneven batch than the three their registers
// would otherwise allow; the smaller classes fit four without spilling.
template <int HM, int KC, bool C>
__global__ void __launch_bounds__(256, (HM >= 12 ? 4 : 1))
    mls_tiled_kernel(const CloudJob* __restrict__ jobs, int njobs, int64_t total) {
  constexpr int LD = CT + 2 * HM;
  constexpr int MWD = 2 * HM + 1;
  __shared__ float4 T[LD][LD + 1];
  __shared__ float Wc[KC > 0 ? KC : 1][256];
  __shared__ double etab[256];  // 2^(j/256) for exp_f32_w (published by the first tile's barrier)
  etab[threadIdx.x] = kExp2Tab256[threadIdx.x];  // 256 threads
  const XcdRange xr = xcd_range(total);
  const int lx = threadIdx.x % CT, ly = threadIdx.x / CT;
  int jprev = 0;
  for (int64_t t = xr.first; t < xr.end; t += xr.stride) {
    const int j = find_job(jobs, njobs, t, jprev);
    jprev = j;
    const CloudJob& cj = jobs[j];
    if ((cj.mls_sem == PSGC_SEM_GPU) != C) continue;  // the same tile for the whole workgroup
    const int64_t lt = t - cj.tile_base;
    const int x0 = (int)(lt % cj.tiles_x) * CT, y0 = (int)(lt / cj.tiles_x) * CT;
    const int hw = cj.mls_hw;
    load_cloud_tile<LD>(cj.mls_src, cj.mls_P, cj.mls_md, cj.mls_thr, cj.rows, cj.cols, y0, x0,
                        hw, T);
    __syncthreads();
    const int x = x0 + lx, y = y0 + ly;
    const bool inimg = x < cj.cols && y < cj.rows;
    const size_t p = (size_t)y * cj.cols + x;
    const int cy = ly + hw, cx = lx + hw;
    const int Wd = 2 * hw + 1;
    const float inv2s2 = cj.mls_inv2s2, r2 = cj.mls_r2;
    const float4 q = T[cy][cx];
    const float pi[3] = {q.x, q.y, q.z};
    float out[3] = {pi[0], pi[1], pi[2]};
    // bits[dy] bit dx: window entry (dy - hw, dx - hw) is a valid point
    // within the cutoff radius (the d^2 test of the CPU restatement, so the
    // neighbour set and every order fallback decision are identical). The
    // three accumulation passes keep the filter's arithmetic
    // (JSLAB_MLS_Filter.cpp:196-340, .cu:214-343): weighted centroid,
    // covariance about it, then the normal equations, with the same
    // contraction as the restatement; an ill-conditioned fit amplifies any
    // reassociation far beyond the 1e-3 mm output tolerance.
    uint32_t bits[MWD];
#pragma unroll
    for (int w = 0; w < MWD; ++w) bits[w] = 0u;
    MlsFit f;
    f.cx = f.cy = f.cz = 0.f;
    float wsum = 0.f;
    int nv = 0;
    int rows_fit = 0;  // leading window rows whose weights all sit in the cache
    bool fit = false, nan_in = false;
    bool mine = false;
    if (inimg) {
      mine = q.w != 0.f;
      if (cj.mls_Pout) st_global(cj.mls_Pout, p, (uint8_t)(mine ? 1 : 0));
      // points removed by SOR enter MLS as zeros (the SOR output), and MLS
      // copies its input for points outside the mask
      if (!mine && cj.mls_md) out[0] = out[1] = out[2] = 0.f;
    }
    // The window scan runs for the whole wave when any of its points is
    // valid (the row loop is then uniform and the entries of a row unroll);
    // an entry is a neighbour when it is valid and not farther than the
    // cutoff, so a NaN distance counts, as in the reference's test.
    if (__any(mine)) {
      for (int dy = 0; dy < Wd; ++dy) {
        const int ry = cy - hw + dy;
        uint32_t row_bits = 0u;
#pragma unroll
        for (int dx = 0; dx < MWD; ++dx) {
          if (dx < Wd) {
            // one 16-byte read: point and mask (an invalid entry's distance
            // is computed and discarded)
            const float4 e4 = T[ry][cx - hw + dx];
            const float pk[3] = {e4.x, e4.y, e4.z};
            const float d2 = dist2<C>(pi[0], pi[1], pi[2], pk);
            if (mine & (e4.w != 0.f) & !(d2 > r2)) {
              row_bits |= 1u << dx;
              const float w = exp_f32_w(-d2 * inv2s2, etab);
              if (KC > 0 && nv < KC) Wc[nv][threadIdx.x] = w;
              mls_accum_centroid<C>(pk, w, &f, &wsum);
              nv++;
            }
          }
        }
        bits[dy] = row_bits;
        if (nv <= KC) rows_fit = dy + 1;
      }
      // A NaN coordinate among the neighbours (or of the point itself)
      // reaches the centroid sums; the reference's fit then yields NaN in
      // every coordinate (its weight, centroid and frame are NaN), which is
      // written directly. Weights here are never NaN (a NaN distance gives
      // the weight of the clamped argument), so the sums of such a point
      // are otherwise meaningless.
      nan_in = (f.cx != f.cx) | (f.cy != f.cy) | (f.cz != f.cz);
      fit = mine && nv >= 3 && (nan_in || !(wsum < 1e-12f));
    }
    // rows served from the cache: the same for every lane of the wave (every
    // lane is active here), so the passes below do not diverge on it
    int dyc = 0;
    if (KC > 0) dyc = __builtin_amdgcn_readfirstlane(wave_min(fit ? rows_fit : MWD));
    if (fit && nan_in) {
      out[0] = out[1] = out[2] = __builtin_nanf("");
    } else if (fit) {
      const float iw = 1.0f / wsum;
      f.cx *= iw;
      f.cy *= iw;
      f.cz *= iw;
      float cov[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
      mls_entries<LD, C>(bits, dyc, Wd, cy - hw, cx - hw, T, Wc, etab, pi, inv2s2,
                         [&](const float pk[3], float w) { mls_accum_cov<C>(f, pk, w, cov); });
      mls_frame(cov, &f);
      const int cur = mls_order(cj.mls_order, nv, cj.mls_sem);
      if (cur == 0) {
        mls_plane(f, pi, out);
      } else if (cj.mls_order >= 2) {  // uniform per tile
        float A[6][6] = {}, bv[6] = {};
        mls_entries<LD, C>(bits, dyc, Wd, cy - hw, cx - hw, T, Wc, etab, pi, inv2s2,
                           [&](const float pk[3], float w) { mls_accum_fixed<6, C>(f, pk, w, A, bv); });
        mls_solve_eval_fixed<6>(f, A, bv, cur, pi, out);
      } else {
        float A[3][3] = {}, bv[3] = {};
        mls_entries<LD, C>(bits, dyc, Wd, cy - hw, cx - hw, T, Wc, etab, pi, inv2s2,
                           [&](const float pk[3], float w) { mls_accum_fixed<3, C>(f, pk, w, A, bv); });
        mls_solve_eval_fixed<3>(f, A, bv, cur, pi, out);
      }
    }
    if (inimg) {
      st_global(cj.mls_dst, p * 3 + 0, out[0]);
      st_global(cj.mls_dst, p * 3 + 1, out[1]);
      st_global(cj.mls_dst, p * 3 + 2, out[2]);
    }
    __syncthreads();
  }
}

int round_grid(int64_t total, int grid) {
  int64_t g = grid;
  if (g > total) g = total;
  if (g < 1) g = 1;
  g = (g + kNumXcd - 1) / kNumXcd * kNumXcd;
  return (int)g;
}

}  // namespace

size_t match_tiled_lds(int cols, int knum) {
  // orders beyond 16 bits: no tiled row (the plan runs the plain match)
  if (knum > 32767) return (size_t)1 << 40;
  return (size_t)cols * sizeof(double) + (size_t)((cols + 1) & ~1) * sizeof(int16_t) +
         (size_t)(knum + 1) * 2 * sizeof(int32_t);
}

int match_tiled_max_lds() { return 64 * 1024; }

int launch_decode_tiled(const DecodeJob* jobs, int njobs, int64_t total, int grid,
                        int tile_rows, hipStream_t s) {
  if (total <= 0) return PSGC_OK;
  if (tile_rows != 2 && tile_rows != 4)
    PSGC_FAIL(PSGC_E_INVALID, "decode tile of %d rows (2 or 4)", tile_rows);
  PSGC_PROF_BEGIN("decode", s);
  if (tile_rows == 2)
    decode_tiled_kernel<2><<<round_grid(total, grid), DC, 0, s>>>(jobs, njobs, total);
  else
    decode_tiled_kernel<4><<<round_grid(total, grid), DC, 0, s>>>(jobs, njobs, total);
  PSGC_LAUNCHED();
  PSGC_PROF_END("decode", s);
  return PSGC_OK;
}

int launch_match_tiled(const MatchJob* jobs, int njobs, int64_t total, int grid, int max_cols,
                       int max_knum, int threads, hipStream_t s) {
  if (total <= 0) return PSGC_OK;
  const size_t lds = match_tiled_lds(max_cols, max_knum);
  if (lds > (size_t)match_tiled_max_lds())
    PSGC_FAIL(PSGC_E_UNSUPPORTED, "row match needs %zu B of LDS", lds);
  if (threads != 128 && threads != 256 && threads != 512)
    PSGC_FAIL(PSGC_E_INVALID, "row match tile of %d threads (128, 256 or 512)", threads);
  PSGC_PROF_BEGIN("match", s);
  const int g = round_grid(total, grid);
  if (threads == 128) match_tiled_kernel<128><<<g, 128, lds, s>>>(jobs, njobs, total);
  else if (threads == 512) match_tiled_kernel<512><<<g, 512, lds, s>>>(jobs, njobs, total);
  else match_tiled_kernel<256><<<g, 256, lds, s>>>(jobs, njobs, total);
  PSGC_LAUNCHED();
  PSGC_PROF_END("match", s);
  return PSGC_OK;
}

int launch_median_tiled(const CloudJob* jobs, int njobs, int64_t total, int grid, hipStream_t s) {
  if (total <= 0) return PSGC_OK;
  PSGC_PROF_BEGIN("median5", s);
  median_tiled_kernel<<<round_grid(total, grid), 256, 0, s>>>(jobs, njobs, total);
  PSGC_LAUNCHED();
  PSGC_PROF_END("median5", s);
  return PSGC_OK;
}

int launch_repro_gauss_tiled(const CloudJob* jobs, int njobs, int64_t total, int grid,
                             int max_half, hipStream_t s) {
  if (total <= 0) return PSGC_OK;
  if (max_half > GH_MAX) PSGC_FAIL(PSGC_E_INVALID, "Gaussian window above 7");
  PSGC_PROF_BEGIN("repro_gauss", s);
  repro_gauss_tiled_kernel<<<round_grid(total, grid), 256, 0, s>>>(jobs, njobs, total);
  PSGC_LAUNCHED();
  PSGC_PROF_END("repro_gauss", s);
  return PSGC_OK;
}

int launch_sor_knn_tiled(const CloudJob* jobs, int njobs, int64_t total, int grid, int max_half,
                         int min_k, int max_k, hipStream_t s) {
  if (total <= 0) return PSGC_OK;
  if (max_half > SH_MAX) PSGC_FAIL(PSGC_E_UNSUPPORTED, "SOR half window above %d", SH_MAX);
  const int g = round_grid(total, grid);
  PSGC_PROF_BEGIN("sor_knn", s);
  if (min_k == max_k && max_k == 30) sor_knn_tiled_kernel<30, 30><<<g, 256, 0, s>>>(jobs, njobs, total);
  else if (min_k == max_k && max_k == 20) sor_knn_tiled_kernel<20, 20><<<g, 256, 0, s>>>(jobs, njobs, total);
  else if (max_k <= 8) sor_knn_tiled_kernel<8, 0><<<g, 256, 0, s>>>(jobs, njobs, total);
  else if (max_k <= 16) sor_knn_tiled_kernel<16, 0><<<g, 256, 0, s>>>(jobs, njobs, total);
  else if (max_k <= 32) sor_knn_tiled_kernel<32, 0><<<g, 256, 0, s>>>(jobs, njobs, total);
  else if (max_k <= 64) sor_knn_tiled_kernel<64, 0><<<g, 256, 0, s>>>(jobs, njobs, total);
  else PSGC_FAIL(PSGC_E_INVALID, "sor k=%d above 64", max_k);
  PSGC_LAUNCHED();
  PSGC_PROF_END("sor_knn", s);
  return PSGC_OK;
}

namespace {

template <int HM>
int launch_mls_class(const CloudJob* jobs, int njobs, int64_t total, int g, int variant, int sems,
                     hipStream_t s) {
  constexpr int KC = mls_cache_entries<HM>();
  const bool cached = KC > 0 && variant == kMlsCached;
  // one launch per semantics present (each skips the other's tiles)
  if (sems & kSemBitCpu) {
    if (cached) mls_tiled_kernel<HM, KC, false><<<g, 256, 0, s>>>(jobs, njobs, total);
    else mls_tiled_kernel<HM, 0, false><<<g, 256, 0, s>>>(jobs, njobs, total);
    PSGC_LAUNCHED();
  }
  if (sems & kSemBitGpu) {
    if (cached) mls_tiled_kernel<HM, KC, true><<<g, 256, 0, s>>>(jobs, njobs, total);
    else mls_tiled_kernel<HM, 0, true><<<g, 256, 0, s>>>(jobs, njobs, total);
    PSGC_LAUNCHED();
  }
  return PSGC_OK;
}

}  // namespace

int launch_mls_tiled(const CloudJob* jobs, int njobs, int64_t total, int grid, int cls,
                     int variant, int sems, hipStream_t s) {
  if (total <= 0) return PSGC_OK;
  if (cls < 0 || cls >= kMlsClasses) PSGC_FAIL(PSGC_E_INVALID, "MLS window class %d", cls);
  if (variant != kMlsPlain && variant != kMlsCached)
    PSGC_FAIL(PSGC_E_INVALID, "MLS variant %d", variant);
  PSGC_PROF_BEGIN("mls", s);
  const int g = round_grid(total, grid);
  switch (cls) {
    case 0: PSGC_TRY(launch_mls_class<kMlsClassHalf[0]>(jobs, njobs, total, g, variant, sems, s)); break;
    case 1: PSGC_TRY(launch_mls_class<kMlsClassHalf[1]>(jobs, njobs, total, g, variant, sems, s)); break;
    case 2: PSGC_TRY(launch_mls_class<kMlsClassHalf[2]>(jobs, njobs, total, g, variant, sems, s)); break;
    case 3: PSGC_TRY(launch_mls_class<kMlsClassHalf[3]>(jobs, njobs, total, g, variant, sems, s)); break;
    default: PSGC_TRY(launch_mls_class<kMlsClassHalf[4]>(jobs, njobs, total, g, variant, sems, s));
  }
  static_assert(kMlsClasses == 5, "one case per class");
  PSGC_PROF_END("mls", s);
  return PSGC_OK;
}

}  // namespace psgc
