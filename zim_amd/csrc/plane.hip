// Plane-fit objectives of the calibration refinement, fused: masked moments
// (count, mean, scatter matrix about the mean), the sum of squared residuals
// to a given plane, and the gradient of any objective that depends on the
// points through those moments, for many views in one launch each.
//
// Every sum is FP64 in a fixed order: each workgroup reduces its chunk of
// one view (each thread its points in order, then a fixed tree), and one
// workgroup per view adds the chunks' partial sums in order, so a result does
// not depend on scheduling. A point counts where its selection byte is 1 (a
// boolean mask, or the triangulation's conditioning codes, 1 = conditioned).
#include "psgc_device.h"
#include "psgc_internal.h"

namespace psgc {
namespace {

constexpr int kPT = 256;              // threads per workgroup
constexpr int kPPT = 16;              // points per thread
constexpr int kChunk = kPT * kPPT;    // points per workgroup
constexpr int kMaxSums = 6;

struct PlaneDev {
  const double* xyz;
  const uint8_t* sel;
  int64_t n;
  double* grad;
  int64_t chunk_base;   // first chunk of this view in the launch
  double mu[3];
  double a[9];          // backward: grad = a (x - mu) + b
  double b[3];
  double nrm[3];        // residuals: plane normal
};

__device__ inline int find_view(const PlaneDev* v, int nv, int64_t chunk) {
  int lo = 0, hi = nv - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (v[mid].chunk_base <= chunk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Fixed-order reduction of NS sums over the workgroup; thread 0 returns them.
template <int NS>
__device__ inline void block_reduce(double (&s)[NS], double (*sh)[kPT]) {
#pragma unroll
  for (int k = 0; k < NS; ++k) sh[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = kPT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int k = 0; k < NS; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) s[k] = sh[k][0];
}

// MODE 0: count and the sum of the points; 1: the six products of the
// centred coordinates (xx, xy, xz, yy, yz, zz); 2: sum of squared residuals
// to the plane through mu with normal nrm, and the sum of the residuals.
template <int MODE>
__global__ void __launch_bounds__(kPT) plane_sums_kernel(const PlaneDev* __restrict__ views,
                                                         int nv, double* __restrict__ part) {
  constexpr int NS = MODE == 0 ? 4 : (MODE == 1 ? 6 : 2);
  __shared__ double sh[NS][kPT];
  const int64_t chunk = blockIdx.x;
  const PlaneDev& v = views[find_view(views, nv, chunk)];
  const int64_t i0 = (chunk - v.chunk_base) * kChunk;
  double s[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) s[k] = 0.0;
  for (int j = 0; j < kPPT; ++j) {
    const int64_t i = i0 + threadIdx.x + (int64_t)j * kPT;
    if (i >= v.n || ld_global(v.sel, i) != 1) continue;
    const double x = ld_global(v.xyz, 3 * i), y = ld_global(v.xyz, 3 * i + 1),
                 z = ld_global(v.xyz, 3 * i + 2);
    if constexpr (MODE == 0) {
      s[0] += 1.0;
      s[1] += x;
      s[2] += y;
      s[3] += z;
    } else {
      const double dx = x - v.mu[0], dy = y - v.mu[1], dz = z - v.mu[2];
      if constexpr (MODE == 1) {
        s[0] += dx * dx;
        s[1] += dx * dy;
        s[2] += dx * dz;
        s[3] += dy * dy;
        s[4] += dy * dz;
        s[5] += dz * dz;
      } else {
        const double r = dx * v.nrm[0] + dy * v.nrm[1] + dz * v.nrm[2];
        s[0] += r * r;
        s[1] += r;
      }
    }
  }
  block_reduce<NS>(s, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k) st_global(part, chunk * kMaxSums + k, s[k]);
  }
}

// One workgroup per view adds its chunks' partial sums (thread t: chunks t,
// t + kPT, ... in order, then the fixed tree) into out[view * NS + k].
template <int NS>
__global__ void __launch_bounds__(kPT) plane_finish_kernel(const PlaneDev* __restrict__ views,
                                                           const double* __restrict__ part,
                                                           double* __restrict__ out) {
  __shared__ double sh[NS][kPT];
  const PlaneDev& v = views[blockIdx.x];
  const int64_t nc = (v.n + kChunk - 1) / kChunk;
  double s[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) s[k] = 0.0;
  for (int64_t c = threadIdx.x; c < nc; c += kPT) {
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] += ld_global(part, (v.chunk_base + c) * kMaxSums + k);
  }
  block_reduce<NS>(s, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k) st_global(out, (int64_t)blockIdx.x * NS + k, s[k]);
  }
}

__global__ void __launch_bounds__(kPT) plane_grad_kernel(const PlaneDev* __restrict__ views,
                                                         int nv) {
  const int64_t chunk = blockIdx.x;
  const PlaneDev& v = views[find_view(views, nv, chunk)];
  const int64_t i0 = (chunk - v.chunk_base) * kChunk;
  for (int j = 0; j < kPPT; ++j) {
    const int64_t i = i0 + threadIdx.x + (int64_t)j * kPT;
    if (i >= v.n) break;
    double g[3] = {0.0, 0.0, 0.0};
    if (ld_global(v.sel, i) == 1) {
      const double dx = ld_global(v.xyz, 3 * i) - v.mu[0];
      const double dy = ld_global(v.xyz, 3 * i + 1) - v.mu[1];
      const double dz = ld_global(v.xyz, 3 * i + 2) - v.mu[2];
#pragma unroll
      for (int r = 0; r < 3; ++r)
        g[r] = v.a[3 * r] * dx + v.a[3 * r + 1] * dy + v.a[3 * r + 2] * dz + v.b[r];
    }
    st_global(v.grad, 3 * i, g[0]);
    st_global(v.grad, 3 * i + 1, g[1]);
    st_global(v.grad, 3 * i + 2, g[2]);
  }
}

// Host-side view table; returns the number of chunks of the launch.
int plane_table(psgc_ctx* ctx, const psgc_plane_view* views, int n, bool need_grad,
                std::vector<PlaneDev>* hv, int64_t* chunks) {
  if (n < 1) PSGC_FAIL(PSGC_E_INVALID, "no views");
  hv->assign(n, PlaneDev{});
  int64_t c = 0;
  for (int v = 0; v < n; ++v) {
    const psgc_plane_view& w = views[v];
    if (w.n < 0 || w.n > ((int64_t)1 << 40)) PSGC_FAIL(PSGC_E_INVALID, "view %d: %lld points", v,
                                                       (long long)w.n);
    if (w.n > 0 && (!w.xyz || !w.sel)) PSGC_FAIL(PSGC_E_INVALID, "view %d: null xyz or sel", v);
    if (need_grad && w.n > 0 && !w.grad_xyz)
      PSGC_FAIL(PSGC_E_INVALID, "view %d: null grad_xyz", v);
    PlaneDev& d = (*hv)[v];
    d.xyz = w.xyz;
    d.sel = w.sel;
    d.n = w.n;
    d.grad = w.grad_xyz;
    d.chunk_base = c;
    c += (w.n + kChunk - 1) / kChunk;
  }
  *chunks = c;
  return PSGC_OK;
}

int plane_upload(psgc_ctx* ctx, const std::vector<PlaneDev>& hv, PlaneDev** d) {
  void* p;
  PSGC_TRY(ctx_buf(ctx, S_PLANEV, sizeof(PlaneDev) * hv.size(), &p));
  PSGC_HIP(hipMemcpyAsync(p, hv.data(), sizeof(PlaneDev) * hv.size(), hipMemcpyHostToDevice,
                          ctx->stream));
  *d = (PlaneDev*)p;
  return PSGC_OK;
}

// One summing pass (MODE) over every view, then the per-view finish; the
// NS sums of each view are returned in host (view * NS + k). The call
// returns when they are there (and the view table's host copy is done).
template <int MODE>
int plane_pass(psgc_ctx* ctx, const std::vector<PlaneDev>& hv, int64_t chunks,
               std::vector<double>* host) {
  constexpr int NS = MODE == 0 ? 4 : (MODE == 1 ? 6 : 2);
  hipStream_t s = ctx->stream;
  PlaneDev* d;
  PSGC_TRY(plane_upload(ctx, hv, &d));
  void *part, *out;
  PSGC_TRY(ctx_buf(ctx, S_PLANEP, sizeof(double) * kMaxSums * (size_t)(chunks > 0 ? chunks : 1),
                   &part));
  PSGC_TRY(ctx_buf(ctx, S_PLANEO, sizeof(double) * NS * hv.size(), &out));
  if (chunks > 0) {
    plane_sums_kernel<MODE><<<(unsigned)chunks, kPT, 0, s>>>(d, (int)hv.size(), (double*)part);
    PSGC_LAUNCHED();
  }
  plane_finish_kernel<NS><<<(unsigned)hv.size(), kPT, 0, s>>>(d, (const double*)part,
                                                              (double*)out);
  PSGC_LAUNCHED();
  host->assign((size_t)NS * hv.size(), 0.0);
  PSGC_HIP(hipMemcpyAsync(host->data(), out, sizeof(double) * NS * hv.size(),
                          hipMemcpyDeviceToHost, s));
  PSGC_HIP(hipStreamSynchronize(s));
  return PSGC_OK;
}

}  // namespace
}  // namespace psgc

using namespace psgc;

extern "C" {

int psgc_plane_moments_batch(psgc_ctx* ctx, const psgc_plane_view* views, int n, double* count,
                             double* mean, double* scatter) {
  PSGC_CTX_ENTER(ctx);
  if (!views || !count || !mean || !scatter) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  std::vector<PlaneDev> hv;
  int64_t chunks;
  PSGC_TRY(plane_table(ctx, views, n, false, &hv, &chunks));
  std::vector<double> s1, s2;
  PSGC_PROF_BEGIN("plane_moments", ctx->stream);
  // pass 1: count and sums -> the mean, which pass 2 centres on
  PSGC_TRY(plane_pass<0>(ctx, hv, chunks, &s1));
  for (int v = 0; v < n; ++v) {
    count[v] = s1[4 * v];
    for (int k = 0; k < 3; ++k) {
      hv[v].mu[k] = s1[4 * v] > 0 ? s1[4 * v + 1 + k] / s1[4 * v] : 0.0;
      mean[3 * v + k] = hv[v].mu[k];
    }
  }
  PSGC_TRY(plane_pass<1>(ctx, hv, chunks, &s2));
  PSGC_PROF_END("plane_moments", ctx->stream);
  for (int v = 0; v < n; ++v) {
    const double* q = &s2[6 * v];
    const double m[9] = {q[0], q[1], q[2], q[1], q[3], q[4], q[2], q[4], q[5]};
    for (int k = 0; k < 9; ++k) scatter[9 * v + k] = m[k];
  }
  return PSGC_OK;
}

int psgc_plane_residuals_batch(psgc_ctx* ctx, const psgc_plane_view* views, int n,
                               const double* mean, const double* normal, double* sumsq,
                               double* sum) {
  PSGC_CTX_ENTER(ctx);
  if (!views || !mean || !normal || !sumsq || !sum) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  std::vector<PlaneDev> hv;
  int64_t chunks;
  PSGC_TRY(plane_table(ctx, views, n, false, &hv, &chunks));
  for (int v = 0; v < n; ++v)
    for (int k = 0; k < 3; ++k) {
      hv[v].mu[k] = mean[3 * v + k];
      hv[v].nrm[k] = normal[3 * v + k];
    }
  std::vector<double> s2;
  PSGC_PROF_BEGIN("plane_residuals", ctx->stream);
  PSGC_TRY(plane_pass<2>(ctx, hv, chunks, &s2));
  PSGC_PROF_END("plane_residuals", ctx->stream);
  for (int v = 0; v < n; ++v) {
    sumsq[v] = s2[2 * v];
    sum[v] = s2[2 * v + 1];
  }
  return PSGC_OK;
}

int psgc_plane_grad_batch(psgc_ctx* ctx, const psgc_plane_view* views, int n, const double* mean,
                          const double* a, const double* b) {
  PSGC_CTX_ENTER(ctx);
  if (!views || !mean || !a || !b) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  std::vector<PlaneDev> hv;
  int64_t chunks;
  PSGC_TRY(plane_table(ctx, views, n, true, &hv, &chunks));
  for (int v = 0; v < n; ++v) {
    for (int k = 0; k < 3; ++k) {
      hv[v].mu[k] = mean[3 * v + k];
      hv[v].b[k] = b[3 * v + k];
    }
    for (int k = 0; k < 9; ++k) hv[v].a[k] = a[9 * v + k];
  }
  PlaneDev* d;
  PSGC_TRY(plane_upload(ctx, hv, &d));
  PSGC_PROF_BEGIN("plane_grad", ctx->stream);
  if (chunks > 0) {
    plane_grad_kernel<<<(unsigned)chunks, kPT, 0, ctx->stream>>>(d, n);
    PSGC_LAUNCHED();
  }
  PSGC_PROF_END("plane_grad", ctx->stream);
  // the view table's host copy ends with this call
  PSGC_HIP(hipStreamSynchronize(ctx->stream));
  return PSGC_OK;
}

}  // extern "C"
