// pipeline.hip — the batched reconstruction pipeline (psgc_batch_*) and the
// tiled single-pair paths of psgc_decode / psgc_reconstruct.
//
// A plan holds, per stage, one table of work descriptors covering every pair
// of the batch, the intermediates of every pair in one device arena, and the
// persistent grid size of every stage. Running the plan issues one launch
// per stage for the whole batch (plus a few tiny per-pair statistics
// kernels), so a captured hipGraph replays the full workload with one call.

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "psgc_internal.h"
#include "psgc_tiled.h"

namespace psgc {

namespace {

constexpr size_t kAlign = 256;

struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
  void* take(size_t bytes) {
    used = (used + kAlign - 1) / kAlign * kAlign;
    void* p = base + used;
    used += bytes;
    return p;
  }
};

// Per-pair operations that run as small per-pair kernels after the batched
// stages (statistics and mask application).
struct PairPost {
  int outlier;              // PSGC_OUTLIER_*
  int use_mls;
  int rows, cols;
  // ROR
  const float* ror_src;
  const uint8_t* ror_P;
  float* ror_dst;
  uint8_t* ror_Pdst;
  float radius;
  int min_nb, window;
  // SOR statistics
  const float* sor_src;
  double* md;
  double* partial;
  double* stats;
  float std_mul;
  float* sor_dst;           // when no MLS (or a one-thread-per-pixel MLS): filtered xyz
  uint8_t* sor_Pdst;        // ... and its mask
  // SOR window beyond the tiled kernel: the one-thread-per-pixel kNN
  int sor_plain, sor_k, sor_half;
  const uint8_t* sor_P;
  // MLS window beyond the tiled kernel: one-thread-per-pixel iterations
  int mls_plain, mls_hw, mls_order, mls_sem, mls_iters;
  float mls_sigma, mls_r2;
  const float* mls_src;
  const uint8_t* mls_P;
  float* mls_tmp;
  float* mls_dst;
  // copies when neither an outlier stage writes the outputs nor MLS runs
  const float* copy_xyz_src;
  const uint8_t* copy_P_src;
  float* copy_xyz_dst;
  uint8_t* copy_P_dst;
  // disparity output copy (no median)
  const float* disp_src;
  float* disp_dst;
};

// A camera decoded by the one-thread-per-pixel kernel (the HDR fit).
struct PlainDecode {
  DecodeArgs a;
  double* phase;
  int32_t* k;
  double* mod;
  uint8_t* valid;
};

// Validity-weighted post-smooth of an HDR camera.
struct HdrSmooth {
  const double* in;
  const uint8_t* valid;
  double* out;
  int rows, cols;
};

// A pair whose rows exceed the tiled match's LDS budget: post-smooth (when
// the decode left it to the match) and the one-thread-per-pixel match.
struct PlainMatch {
  const double* lp_in;
  const double* rp_in;
  double* lp_smooth;  // NULL: lp_in is already smoothed
  double* rp_smooth;
  const int32_t* lk;
  const int32_t* rk;
  int rows, cols, pha_f;
  double grad_thr;
  int32_t* span;
  double* disp64;
  float* disp32;
  int32_t* match_col;
};

// Host data copied into the arena once the plan is built.
struct Blob {
  char* dev;
  std::vector<char> host;
};

template <typename J>
struct Table {
  std::vector<J> host;
  J* dev = nullptr;
  int64_t total = 0;
};

// Semantics bits of a table of MLS jobs.
int mls_sems(const std::vector<CloudJob>& jobs) {
  int m = 0;
  for (const CloudJob& j : jobs) m |= mls_sem_bit(j.mls_sem);
  return m;
}

}  // namespace

// Row-match jobs of one image width: launched on their own, the LDS row
// buffer is sized for that width (narrow rows then keep more workgroups per
// CU), with a grid and threads per row of their own.
struct MatchGroup {
  int cols = 0, knum = 0;
  int grid = 0;        // 0: the batch-wide "match" grid
  int threads = 256;
  Table<MatchJob> tab;
};

struct Plan {
  psgc_ctx* ctx = nullptr;
  int n = 0;
  Arena arena;
  Table<DecodeJob> dec;
  Table<MatchJob> mat;
  std::vector<MatchGroup> match_groups;  // by width, when the batch has several
  bool match_grouped = false;            // one launch per width instead of one
  Table<CloudJob> med, rg, sor;
  Table<SorStatJob> sst;             // SOR statistics, one job per pair
  // MLS jobs by iteration and window class (one launch each)
  std::vector<std::array<Table<CloudJob>, kMlsClasses>> mls;
  std::vector<PairPost> post;
  std::vector<DecodeArgs> gray_fix;       // gray-code jump correction per camera
  std::vector<const double*> gray_in_ph;
  std::vector<double*> gray_out_ph;
  std::vector<const int32_t*> gray_in_k;
  std::vector<int32_t*> gray_out_k;
  std::vector<const uint8_t*> gray_valid;
  std::vector<PlainDecode> plain_dec;
  std::vector<HdrSmooth> hdr_smooth;
  std::vector<PlainMatch> plain_match;
  std::vector<Blob> blobs;
  int max_cols = 0, max_knum = 0, max_ghalf = 0, max_sor_half = 0, max_sor_k = 0, max_mls_hw = 0;
  int min_sor_k = 1 << 30;
  std::map<std::string, int> grid;  // per stage; "decode_rows": the decode tile height
  int dec_rows = 4;
  int match_threads = 256;
  int mls_variant = kMlsPlain;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~Plan() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    if (arena.base) hipFree(arena.base);
  }
};

namespace {

int tiles(int n, int t) { return (n + t - 1) / t; }

// Lays the decode tiles out as kDecodeTileCols x tile_rows per job.
void layout_decode(Plan* pl, int tile_rows) {
  pl->dec_rows = tile_rows;
  pl->dec.total = 0;
  for (DecodeJob& j : pl->dec.host) {
    j.tiles_x = tiles(j.a.cols, kDecodeTileCols);
    j.tiles_y = tiles(j.a.rows, tile_rows);
    j.tile_base = pl->dec.total;
    pl->dec.total += (int64_t)j.tiles_x * j.tiles_y;
  }
}

int upload_tables(Plan* pl) {
  // device copies of the tables live at the end of the arena
  auto put = [&](auto& tab) -> int {
    using J = typename std::decay_t<decltype(tab.host)>::value_type;
    if (tab.host.empty()) return PSGC_OK;
    tab.dev = (J*)pl->arena.take(tab.host.size() * sizeof(J));
    PSGC_HIP(hipMemcpy(tab.dev, tab.host.data(), tab.host.size() * sizeof(J),
                       hipMemcpyHostToDevice));
    return PSGC_OK;
  };
  for (const Blob& b : pl->blobs)
    PSGC_HIP(hipMemcpy(b.dev, b.host.data(), b.host.size(), hipMemcpyHostToDevice));
  PSGC_TRY(put(pl->dec));
  PSGC_TRY(put(pl->mat));
  for (MatchGroup& mg : pl->match_groups) PSGC_TRY(put(mg.tab));
  PSGC_TRY(put(pl->med));
  PSGC_TRY(put(pl->rg));
  PSGC_TRY(put(pl->sor));
  PSGC_TRY(put(pl->sst));
  for (auto& it : pl->mls)
    for (auto& t : it) PSGC_TRY(put(t));
  return PSGC_OK;
}

size_t table_bytes(const Plan* pl) {
  size_t b = 0;
  auto add = [&](const auto& tab) {
    using J = typename std::decay_t<decltype(tab.host)>::value_type;
    b += tab.host.size() * sizeof(J) + kAlign;
  };
  add(pl->dec);
  add(pl->mat);
  for (const MatchGroup& mg : pl->match_groups) add(mg.tab);
  add(pl->med);
  add(pl->rg);
  add(pl->sor);
  add(pl->sst);
  for (auto& it : pl->mls)
    for (auto& t : it) add(t);
  return b;
}

// Row tile before any tuning: rows wider than 2048 pixels fill most of a
// workgroup's LDS (12 bytes per pixel), so few workgroups fit a CU and
// wider ones keep more waves resident.
int default_match_threads(int cols) { return cols > 2048 ? 512 : 256; }

// Splits the row-match jobs by image width (each group's rows renumbered
// from 0); a batch of one width keeps the single table only.
void group_matches(Plan* pl) {
  pl->match_threads = default_match_threads(pl->max_cols);
  pl->match_groups.clear();
  std::map<int, size_t> at;
  for (const MatchJob& j : pl->mat.host)
    if (!at.count(j.cols)) {
      at[j.cols] = pl->match_groups.size();
      MatchGroup mg;
      mg.cols = j.cols;
      mg.threads = default_match_threads(j.cols);
      pl->match_groups.push_back(mg);
    }
  if (pl->match_groups.size() < 2) {
    pl->match_groups.clear();
    pl->match_grouped = false;
    return;
  }
  for (MatchJob j : pl->mat.host) {
    MatchGroup& mg = pl->match_groups[at[j.cols]];
    j.tile_base = mg.tab.total;
    mg.tab.total += j.rows;
    mg.knum = std::max(mg.knum, j.pha_f);
    mg.tab.host.push_back(j);
  }
}

int default_grid(psgc_ctx* ctx, int per_cu) {
  const int cu = ctx->cu_count > 0 ? ctx->cu_count : 256;
  return cu * per_cu;
}

}  // namespace

// Input of one pair to the reconstruct part of a plan.
struct ReconInput {
  const double* lph;   // phase maps (raw if smooth)
  const double* rph;
  const int32_t* lk;
  const int32_t* rk;
  int smooth;          // post-smooth still to apply (fused into the match)
};

// Builds the reconstruct stages of pair i (after decode) into the plan.
// Buffers come from the arena (sizes counted by a first dry pass when
// `count` is set: then only arena.used advances).
static int plan_reconstruct(Plan* pl, int i, const psgc_recon_config& rc, int rows, int cols,
                            int pha_f, const double* Q, const float* const* coef,
                            const double* R1, const ReconInput& in,
                            float* xyz_out, uint8_t* P_out, float* disp_out, int32_t* mc_out,
                            double* d64_out) {
  Arena& A = pl->arena;
  const size_t np = (size_t)rows * cols;
  const int gw = rc.gauss_window;
  if (gw < 1 || gw > 7 || (gw & 1) == 0) PSGC_FAIL(PSGC_E_INVALID, "gauss_window=%d", gw);
  if (rc.outlier_mode < PSGC_OUTLIER_NONE || rc.outlier_mode > PSGC_OUTLIER_SOR)
    PSGC_FAIL(PSGC_E_INVALID, "outlier_mode=%d", rc.outlier_mode);
  if (rc.outlier_mode == PSGC_OUTLIER_SOR) {
    if (rc.sor_k < 1 || rc.sor_k > 64) PSGC_FAIL(PSGC_E_INVALID, "sor_k=%d", rc.sor_k);
    if (rc.sor_window < 1 || rc.sor_window > 512)
      PSGC_FAIL(PSGC_E_INVALID, "sor_window=%d", rc.sor_window);
  }
  if (rc.use_mls) {
    PSGC_TRY(check_mls_weights(rc.mls_spatial_sigma, rc.mls_max_radius));
    if (rc.mls_window_size < 0 || rc.mls_window_size > 256)
      PSGC_FAIL(PSGC_E_INVALID, "mls_window_size=%d", rc.mls_window_size);
  }
  // Beyond what the tiled kernels hold (SOR and MLS windows of half size 8
  // and 10, a row and its order spans in the match's LDS budget) a pair's
  // stage runs on the one-thread-per-pixel kernels instead, with the same
  // results.
  const bool sor_plain = rc.outlier_mode == PSGC_OUTLIER_SOR && rc.sor_window > kSorTiledMaxHalf;
  const bool mls_plain = rc.use_mls && rc.mls_window_size > kMlsTiledMaxHalf;
  const bool match_plain = match_tiled_lds(cols, pha_f) > (size_t)match_tiled_max_lds();
  const bool cpu = rc.semantics == PSGC_SEM_CPU;
  const bool use64 = !rc.use_disp_median_filter && cpu;

  // ---- match
  float* disp = disp_out && !rc.use_disp_median_filter ? disp_out : (float*)A.take(np * 4);
  double* d64 = d64_out ? d64_out : (use64 ? (double*)A.take(np * 8) : nullptr);
  if (match_plain) {
    PlainMatch pm{};
    pm.lp_in = in.lph;
    pm.rp_in = in.rph;
    if (in.smooth) {
      pm.lp_smooth = (double*)A.take(np * 8);
      pm.rp_smooth = (double*)A.take(np * 8);
    }
    pm.lk = in.lk;
    pm.rk = in.rk;
    pm.rows = rows;
    pm.cols = cols;
    pm.pha_f = pha_f;
    pm.grad_thr = rc.grad_threshold;
    pm.span = (int32_t*)A.take((size_t)rows * (pha_f + 1) * 2 * sizeof(int32_t));
    pm.disp64 = d64;
    pm.disp32 = disp;
    pm.match_col = mc_out;
    pl->plain_match.push_back(pm);
  } else {
    MatchJob mj{};
    mj.lp = in.lph;
    mj.rp = in.rph;
    mj.lk = in.lk;
    mj.rk = in.rk;
    mj.rows = rows;
    mj.cols = cols;
    mj.pha_f = pha_f;
    mj.smooth = in.smooth;
    mj.grad_thr = rc.grad_threshold;
    mj.disp = disp;
    mj.disp64 = d64;
    mj.match_col = mc_out;
    mj.tile_base = pl->mat.total;
    pl->mat.total += rows;
    pl->mat.host.push_back(mj);
    pl->max_cols = std::max(pl->max_cols, cols);
    pl->max_knum = std::max(pl->max_knum, pha_f);
  }

  CloudJob base{};
  base.rows = rows;
  base.cols = cols;
  base.tiles_x = tiles(cols, 16);
  base.tiles_y = tiles(rows, 16);
  const int64_t ntile = (int64_t)base.tiles_x * base.tiles_y;

  // ---- median
  const float* dsrc = disp;
  if (rc.use_disp_median_filter) {
    CloudJob c = base;
    c.tile_base = pl->med.total;
    pl->med.total += ntile;
    c.disp_in = disp;
    c.disp_med = disp_out ? disp_out : (float*)A.take(np * 4);
    dsrc = c.disp_med;
    pl->med.host.push_back(c);
  }

  // ---- reproject + Gaussian
  float* xyzg = (float*)A.take(np * 12);
  uint8_t* Pg = (uint8_t*)A.take(np);
  uint8_t* P0 = cpu ? (uint8_t*)A.take(np) : nullptr;
  {
    CloudJob c = base;
    c.tile_base = pl->rg.total;
    pl->rg.total += ntile;
    c.disp_src = dsrc;
    c.disp_src64 = use64 ? d64 : nullptr;
    if (coef) {
      for (int a = 0; a < 3; ++a) c.coef[a] = (const float4*)coef[a];
    } else {
      std::memcpy(c.q, Q, sizeof(c.q));
    }
    std::memcpy(c.r1, R1, sizeof(c.r1));
    c.min_depth = rc.min_depth;
    c.max_depth = rc.max_depth;
    c.use_rect = rc.use_rectified_coords;
    c.semantics = rc.semantics;
    c.gwin = gw;
    gauss_kernel_f32(gw, rc.gauss_sigma, c.gk);
    c.xyz0 = nullptr;
    c.P0 = P0;
    c.xyzg = xyzg;
    c.Pg = Pg;
    pl->rg.host.push_back(c);
    pl->max_ghalf = std::max(pl->max_ghalf, (gw - 1) / 2);
  }
  // CPU semantics: the outlier stage reads the pre-Gaussian mask.
  const uint8_t* Pin = cpu ? P0 : Pg;

  PairPost pp{};
  pp.outlier = rc.outlier_mode;
  pp.use_mls = rc.use_mls;
  pp.rows = rows;
  pp.cols = cols;
  if (disp_out && !rc.use_disp_median_filter) pp.disp_dst = nullptr;  // written by the match

  const float* mls_src = xyzg;
  const uint8_t* mls_mask = Pin;
  const double* mls_md = nullptr;
  const double* mls_thr = nullptr;
  uint8_t* mls_Pout = P_out;
  float* t0 = nullptr;
  float* t1 = rc.use_mls && rc.mls_iterations > 1 ? (float*)A.take(np * 12) : nullptr;

  if (rc.outlier_mode == PSGC_OUTLIER_ROR) {
    float* dst = rc.use_mls ? (t0 = (float*)A.take(np * 12)) : xyz_out;
    pp.ror_src = xyzg;
    pp.ror_P = Pin;
    pp.ror_dst = dst;
    pp.ror_Pdst = P_out;
    pp.radius = rc.radius_search;
    pp.min_nb = rc.radius_min_neighbors;
    pp.window = rc.radius_window_size;
    mls_src = dst;
    mls_mask = P_out;
    mls_Pout = nullptr;
  } else if (rc.outlier_mode == PSGC_OUTLIER_SOR && sor_plain) {
    double* md = (double*)A.take(np * 8);
    double* part = (double*)A.take(sor_partial_doubles(np) * 8);
    double* st = (double*)A.take(4 * 8);
    float* dst = rc.use_mls ? (t0 = (float*)A.take(np * 12)) : xyz_out;
    pp.sor_plain = 1;
    pp.sor_src = xyzg;
    pp.sor_P = Pin;
    pp.sor_k = rc.sor_k;
    pp.sor_half = rc.sor_window;
    pp.md = md;
    pp.partial = part;
    pp.stats = st;
    pp.std_mul = rc.sor_std_mul;
    pp.sor_dst = dst;
    pp.sor_Pdst = P_out;
    mls_src = dst;
    mls_mask = P_out;
    mls_Pout = nullptr;
  } else if (rc.outlier_mode == PSGC_OUTLIER_SOR) {
    double* md = (double*)A.take(np * 8);
    double* part = (double*)A.take(sor_partial_doubles(np) * 8);
    double* st = (double*)A.take(4 * 8);
    CloudJob c = base;
    c.tile_base = pl->sor.total;
    pl->sor.total += ntile;
    c.sor_xyz = xyzg;
    c.sor_P = Pin;
    c.sor_k = rc.sor_k;
    c.sor_half = rc.sor_window;
    c.md = md;
    pl->sor.host.push_back(c);
    pl->max_sor_half = std::max(pl->max_sor_half, rc.sor_window);
    pl->max_sor_k = std::max(pl->max_sor_k, rc.sor_k);
    pl->min_sor_k = std::min(pl->min_sor_k, rc.sor_k);
    pp.sor_src = xyzg;
    pp.md = md;
    pp.partial = part;
    pp.stats = st;
    pp.std_mul = rc.sor_std_mul;
    {
      SorStatJob sj{};
      sj.md = md;
      sj.npix = (size_t)rows * cols;
      sj.nchunks = sor_chunks(sj.npix);
      sj.np2 = sor_partial_doubles(sj.npix) / 3;
      sj.part = part;
      sj.stats = st;
      sj.std_mul = rc.sor_std_mul;
      sj.chunk_base = pl->sst.total;
      pl->sst.total += (int64_t)sj.nchunks;
      pl->sst.host.push_back(sj);
    }
    if (!rc.use_mls) {
      pp.sor_dst = xyz_out;
      pp.sor_Pdst = P_out;
    } else if (mls_plain) {
      // the one-thread-per-pixel MLS does not gate on the statistics itself
      pp.sor_dst = t0 = (float*)A.take(np * 12);
      pp.sor_Pdst = P_out;
      mls_src = t0;
      mls_mask = P_out;
      mls_Pout = nullptr;
    } else {
      mls_md = md;
      mls_thr = st + 3;
    }
  } else if (!rc.use_mls) {
    pp.copy_xyz_src = xyzg;
    pp.copy_P_src = Pin;
    pp.copy_xyz_dst = xyz_out;
    pp.copy_P_dst = P_out;
  } else if (mls_plain) {
    // the mask passes through unchanged
    pp.copy_P_src = Pin;
    pp.copy_P_dst = P_out;
    mls_Pout = nullptr;
  }

  if (rc.use_mls && mls_plain) {
    pp.mls_plain = 1;
    pp.mls_iters = rc.mls_iterations > 1 ? rc.mls_iterations : 1;
    pp.mls_order = std::max(0, std::min(2, rc.mls_polynomial_order));
    pp.mls_r2 = (rc.mls_max_radius > 0.0f) ? rc.mls_max_radius * rc.mls_max_radius
                                           : 9.0f * rc.mls_spatial_sigma * rc.mls_spatial_sigma;
    pp.mls_sigma = rc.mls_spatial_sigma;
    pp.mls_hw = rc.mls_window_size;
    pp.mls_sem = rc.semantics;
    pp.mls_src = mls_src;
    pp.mls_P = mls_mask;
    pp.mls_tmp = t1;
    pp.mls_dst = xyz_out;
  } else if (rc.use_mls) {
    const int n_it = rc.mls_iterations > 1 ? rc.mls_iterations : 1;
    const int order = std::max(0, std::min(2, rc.mls_polynomial_order));
    const float r2 = (rc.mls_max_radius > 0.0f) ? rc.mls_max_radius * rc.mls_max_radius
                                                : 9.0f * rc.mls_spatial_sigma * rc.mls_spatial_sigma;
    if ((int)pl->mls.size() < n_it) pl->mls.resize(n_it);
    const float* src = mls_src;
    for (int it = 0; it < n_it; ++it) {
      CloudJob c = base;
      Table<CloudJob>& tab = pl->mls[it][mls_class(rc.mls_window_size)];
      c.tile_base = tab.total;
      tab.total += ntile;
      float* dst = (it == n_it - 1) ? xyz_out : ((it % 2 == 0) ? t1 : t0);
      if (dst == nullptr || dst == src) dst = (src == t1) ? (t0 ? t0 : (t0 = (float*)A.take(np * 12))) : t1;
      c.mls_src = src;
      c.mls_P = it == 0 ? mls_mask : P_out;
      c.mls_md = it == 0 ? mls_md : nullptr;
      c.mls_thr = it == 0 ? mls_thr : nullptr;
      c.mls_dst = dst;
      c.mls_Pout = it == 0 ? mls_Pout : nullptr;
      c.mls_hw = rc.mls_window_size;
      c.mls_order = order;
      c.mls_sem = rc.semantics;
      c.mls_inv2s2 = 1.0f / (2.0f * rc.mls_spatial_sigma * rc.mls_spatial_sigma);
      c.mls_r2 = r2;
      tab.host.push_back(c);
      src = dst;
    }
    pl->max_mls_hw = std::max(pl->max_mls_hw, rc.mls_window_size);
  }
  pl->post.push_back(pp);
  (void)i;
  return PSGC_OK;
}

static int run_plan(Plan* pl, hipStream_t s) {
  // grid per stage: the autotuner's or a pinned choice, else the workgroups
  // per CU the tuner settles on for full-size pairs (far more than resident:
  // the dispatcher balances uneven tiles)
  auto g = [&](const char* st, int per_cu) {
    auto it = pl->grid.find(st);
    return it != pl->grid.end() ? it->second : default_grid(pl->ctx, per_cu);
  };
  for (const PlainDecode& d : pl->plain_dec)
    PSGC_TRY(launch_decode_plain(d.a, d.phase, d.k, d.mod, d.valid, s));
  if (!pl->dec.host.empty())
    PSGC_TRY(launch_decode_tiled(pl->dec.dev, (int)pl->dec.host.size(), pl->dec.total,
                                 g("decode", 64), pl->dec_rows, s));
  for (size_t c = 0; c < pl->gray_fix.size(); ++c)
    PSGC_TRY(launch_gray_fix(pl->gray_fix[c], pl->gray_in_ph[c], pl->gray_out_ph[c],
                             pl->gray_in_k[c], pl->gray_out_k[c], pl->gray_valid[c], s));
  for (const HdrSmooth& h : pl->hdr_smooth)
    PSGC_TRY(launch_post_smooth_valid(h.in, h.valid, h.out, h.rows, h.cols, s));
  if (pl->match_grouped) {
    for (const MatchGroup& mg : pl->match_groups)
      PSGC_TRY(launch_match_tiled(mg.tab.dev, (int)mg.tab.host.size(), mg.tab.total,
                                  mg.grid ? mg.grid : g("match", 64), mg.cols, mg.knum,
                                  mg.threads, s));
  } else if (!pl->mat.host.empty()) {
    PSGC_TRY(launch_match_tiled(pl->mat.dev, (int)pl->mat.host.size(), pl->mat.total,
                                g("match", 64), pl->max_cols, pl->max_knum, pl->match_threads, s));
  }
  for (const PlainMatch& m : pl->plain_match) {
    const double* lp = m.lp_in;
    const double* rp = m.rp_in;
    if (m.lp_smooth) {
      PSGC_TRY(launch_post_smooth(m.lp_in, m.lp_smooth, m.rows, m.cols, s));
      PSGC_TRY(launch_post_smooth(m.rp_in, m.rp_smooth, m.rows, m.cols, s));
      lp = m.lp_smooth;
      rp = m.rp_smooth;
    }
    PSGC_TRY(launch_match_plain(lp, rp, m.lk, m.rk, m.rows, m.cols, m.pha_f, m.grad_thr, m.span,
                                m.disp64, m.disp32, m.match_col, s));
  }
  if (!pl->med.host.empty())
    PSGC_TRY(launch_median_tiled(pl->med.dev, (int)pl->med.host.size(), pl->med.total,
                                 g("median5", 64), s));
  if (!pl->rg.host.empty())
    PSGC_TRY(launch_repro_gauss_tiled(pl->rg.dev, (int)pl->rg.host.size(), pl->rg.total,
                                      g("repro_gauss", 128), pl->max_ghalf, s));
  if (!pl->sor.host.empty())
    PSGC_TRY(launch_sor_knn_tiled(pl->sor.dev, (int)pl->sor.host.size(), pl->sor.total,
                                  g("sor_knn", 256), pl->max_sor_half, pl->min_sor_k,
                                  pl->max_sor_k, s));
  // every SOR pair's statistics in two launches
  if (!pl->sst.host.empty())
    PSGC_TRY(launch_sor_stats_batch(pl->sst.dev, (int)pl->sst.host.size(), pl->sst.total, s));
  for (const PairPost& pp : pl->post) {
    const size_t np = (size_t)pp.rows * pp.cols;
    if (pp.outlier == PSGC_OUTLIER_ROR) {
      PSGC_TRY(launch_ror(pp.ror_src, pp.ror_P, pp.ror_dst, pp.ror_Pdst, pp.rows, pp.cols,
                          pp.radius, pp.min_nb, pp.window, s));
    } else if (pp.outlier == PSGC_OUTLIER_SOR && pp.sor_plain) {
      PSGC_TRY(launch_sor(pp.sor_src, pp.sor_P, pp.sor_dst, pp.sor_Pdst, pp.rows, pp.cols,
                          pp.sor_k, pp.std_mul, pp.sor_half, pp.md, pp.partial, pp.stats, s));
    } else if (pp.outlier == PSGC_OUTLIER_SOR) {
      // the tiled MLS gates on the statistics itself; otherwise apply them
      if (pp.sor_dst)
        PSGC_TRY(launch_sor_apply(pp.sor_src, pp.md, pp.sor_dst, pp.sor_Pdst, np, pp.stats, s));
    } else {
      if (pp.copy_xyz_dst)
        PSGC_HIP(hipMemcpyAsync(pp.copy_xyz_dst, pp.copy_xyz_src, np * 12,
                                hipMemcpyDeviceToDevice, s));
      if (pp.copy_P_dst)
        PSGC_HIP(hipMemcpyAsync(pp.copy_P_dst, pp.copy_P_src, np, hipMemcpyDeviceToDevice, s));
    }
    if (pp.mls_plain) {
      const float* in = pp.mls_src;
      for (int it = 0; it < pp.mls_iters; ++it) {
        float* out = ((pp.mls_iters - 1 - it) % 2 == 0) ? pp.mls_dst : pp.mls_tmp;
        PSGC_TRY(launch_mls(in, pp.mls_P, out, pp.rows, pp.cols, pp.mls_hw, pp.mls_sigma,
                            pp.mls_order, pp.mls_r2, pp.mls_sem, s));
        in = out;
      }
    }
  }
  for (auto& it : pl->mls)
    for (int c = 0; c < kMlsClasses; ++c)
      if (!it[c].host.empty())
        PSGC_TRY(launch_mls_tiled(it[c].dev, (int)it[c].host.size(), it[c].total, g("mls", 128),
                                  c, pl->mls_variant, mls_sems(it[c].host), s));
  return PSGC_OK;
}

// Re-lays the decode job table for a tile height and uploads it (the stream
// is drained first: no launch in flight reads the table).
static int relayout_decode(Plan* pl, int tile_rows, hipStream_t s) {
  PSGC_HIP(hipStreamSynchronize(s));
  layout_decode(pl, tile_rows);
  PSGC_HIP(hipMemcpy(pl->dec.dev, pl->dec.host.data(), pl->dec.host.size() * sizeof(DecodeJob),
                     hipMemcpyHostToDevice));
  return PSGC_OK;
}

// Times each stage of the plan over candidate persistent grid sizes (and, for
// the decode, tile heights; for the MLS, kernel variants) and keeps the
// fastest; the plan's inputs of every stage must already be valid.
static int autotune(Plan* pl, hipStream_t s) {
  const int cu = pl->ctx->cu_count > 0 ? pl->ctx->cu_count : 256;
  const int cands[] = {1, 2, 4, 8, 16, 32, 64, 128, 256};
  hipEvent_t a, b;
  PSGC_HIP(hipEventCreate(&a));
  PSGC_HIP(hipEventCreate(&b));
  // milliseconds of the second of two runs of `launch` (the first warms caches)
  auto time2 = [&](auto&& launch, float* ms) -> int {
    float t = 0;
    for (int rep = 0; rep < 2; ++rep) {
      PSGC_HIP(hipEventRecord(a, s));
      PSGC_TRY(launch());
      PSGC_HIP(hipEventRecord(b, s));
      PSGC_HIP(hipEventSynchronize(b));
      PSGC_HIP(hipEventElapsedTime(&t, a, b));
    }
    *ms = t;
    return PSGC_OK;
  };
  struct St {
    const char* name;
    bool on;
  };
  const St stages[] = {{"decode", !pl->dec.host.empty()},    {"match", !pl->mat.host.empty()},
                       {"median5", !pl->med.host.empty()},   {"repro_gauss", !pl->rg.host.empty()},
                       {"sor_knn", !pl->sor.host.empty()},   {"mls", !pl->mls.empty()}};
  for (const St& st : stages) {
    if (!st.on) continue;
    const std::string n = st.name;
    // second tuning dimension: decode tile rows / MLS kernel variant
    std::vector<int> shapes{0};
    if (n == "decode") shapes = {4, 2};
    if (n == "match") shapes = {256, 128, 512};
    if (n == "mls") shapes = {kMlsPlain, kMlsCached};
    float best_ms = 1e30f;
    int best_g = 0, best_shape = shapes[0];
    for (int shape : shapes) {
      if (n == "decode") PSGC_TRY(relayout_decode(pl, shape, s));
      for (int pc : cands) {
        const int grid = cu * pc;
        float tot = 0;
        for (int rep = 0; rep < 2; ++rep) {
          PSGC_HIP(hipEventRecord(a, s));
          if (n == "decode")
            PSGC_TRY(launch_decode_tiled(pl->dec.dev, (int)pl->dec.host.size(), pl->dec.total,
                                         grid, shape, s));
          else if (n == "match")
            PSGC_TRY(launch_match_tiled(pl->mat.dev, (int)pl->mat.host.size(), pl->mat.total,
                                        grid, pl->max_cols, pl->max_knum, shape, s));
          else if (n == "median5")
            PSGC_TRY(launch_median_tiled(pl->med.dev, (int)pl->med.host.size(), pl->med.total,
                                         grid, s));
          else if (n == "repro_gauss")
            PSGC_TRY(launch_repro_gauss_tiled(pl->rg.dev, (int)pl->rg.host.size(), pl->rg.total,
                                              grid, pl->max_ghalf, s));
          else if (n == "sor_knn")
            PSGC_TRY(launch_sor_knn_tiled(pl->sor.dev, (int)pl->sor.host.size(), pl->sor.total,
                                          grid, pl->max_sor_half, pl->min_sor_k,
                                          pl->max_sor_k, s));
          else
            for (int c = 0; c < kMlsClasses; ++c) {
              const Table<CloudJob>& t = pl->mls[0][c];
              if (!t.host.empty())
                PSGC_TRY(launch_mls_tiled(t.dev, (int)t.host.size(), t.total, grid, c, shape,
                                          mls_sems(t.host), s));
            }
          PSGC_HIP(hipEventRecord(b, s));
          PSGC_HIP(hipEventSynchronize(b));
          float ms = 0;
          PSGC_HIP(hipEventElapsedTime(&ms, a, b));
          if (rep > 0) tot += ms;  // first repetition warms caches
        }
        if (tot < best_ms) {
          best_ms = tot;
          best_g = grid;
          best_shape = shape;
        }
      }
    }
    pl->grid[n] = best_g;
    if (n == "decode") {
      PSGC_TRY(relayout_decode(pl, best_shape, s));
      pl->grid["decode_rows"] = best_shape;
    } else if (n == "match") {
      pl->match_threads = best_shape;
      pl->grid["match_threads"] = best_shape;
    } else if (n == "mls") {
      pl->mls_variant = best_shape;
      pl->grid["mls_variant"] = best_shape;
    }
  }
  // per image width: each width's rows on their own, over the same grid and
  // threads-per-row candidates; the widths' launches replace the single one
  // when, run back to back, they beat it
  if (!pl->match_groups.empty()) {
    for (MatchGroup& mg : pl->match_groups) {
      float best = 1e30f;
      for (int threads : {256, 128, 512})
        for (int pc : cands) {
          float ms = 0;
          PSGC_TRY(time2([&]() -> int {
            return launch_match_tiled(mg.tab.dev, (int)mg.tab.host.size(), mg.tab.total,
                                      cu * pc, mg.cols, mg.knum, threads, s);
          }, &ms));
          if (ms < best) {
            best = ms;
            mg.grid = cu * pc;
            mg.threads = threads;
          }
        }
    }
    float single = 0, grouped = 0;
    PSGC_TRY(time2([&]() -> int {
      return launch_match_tiled(pl->mat.dev, (int)pl->mat.host.size(), pl->mat.total,
                                pl->grid["match"], pl->max_cols, pl->max_knum,
                                pl->match_threads, s);
    }, &single));
    PSGC_TRY(time2([&]() -> int {
      for (const MatchGroup& mg : pl->match_groups)
        PSGC_TRY(launch_match_tiled(mg.tab.dev, (int)mg.tab.host.size(), mg.tab.total, mg.grid,
                                    mg.cols, mg.knum, mg.threads, s));
      return PSGC_OK;
    }, &grouped));
    pl->match_grouped = grouped < single;
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  // the MLS tuning pass overwrote iteration 0's output only, which every
  // run recomputes first; nothing else was modified
  return PSGC_OK;
}

// Allocates the arena: a dry build counts bytes, the real build takes them.
template <typename Build>
static int build_plan(Plan* pl, Build&& build) {
  Plan dry;
  dry.ctx = pl->ctx;
  dry.arena.base = nullptr;
  PSGC_TRY(build(&dry));
  group_matches(&dry);
  const size_t bytes = dry.arena.used + table_bytes(&dry) + 4 * kAlign;
  if (hipMalloc(&pl->arena.base, bytes) != hipSuccess) {
    (void)hipGetLastError();
    pl->arena.base = nullptr;
    PSGC_FAIL(PSGC_E_ALLOC, "batch arena of %zu bytes", bytes);
  }
  pl->arena.size = bytes;
  PSGC_TRY(build(pl));
  group_matches(pl);
  return upload_tables(pl);
}

// Decode stages of one camera into the plan; returns the phase/k the match
// should read and whether the post-smooth is still to apply (fused into the
// match).
static int plan_decode(Plan* pl, const psgc_decode_config& dc, const void* const* frames,
                       int n_frames, int sample_type, int rows, int cols, int64_t stride,
                       double* phase_raw, int32_t* k, double* mod, uint8_t* valid,
                       const double** ph_for_match, const int32_t** k_for_match,
                       int* smooth_in_match) {
  DecodeJob j{};
  PSGC_TRY(build_decode_args(&dc, n_frames, sample_type, rows, cols, stride, &j.a));
  for (int f = 0; f < n_frames; ++f) {
    if (!frames[f]) PSGC_FAIL(PSGC_E_INVALID, "frame %d is null", f);
    j.a.frames[f] = frames[f];
  }
  const size_t np = (size_t)rows * cols;
  const bool gray = j.a.nlev == 1;
  const bool hdr = j.a.hdr != 0;
  j.phase = phase_raw;
  j.k = k;
  j.mod = mod;
  j.valid = valid;
  if ((gray || hdr) && !valid) j.valid = (uint8_t*)pl->arena.take(np);
  *ph_for_match = phase_raw;
  *k_for_match = k;
  *smooth_in_match = dc.post_smooth ? 1 : 0;
  if (hdr) {
    // the fit needs a level's raw samples together: one thread per pixel
    Blob b;
    b.host.resize(hdr_table_bytes(j.a));
    hdr_table_fill(j.a, b.host.data());
    b.dev = (char*)pl->arena.take(b.host.size());
    hdr_table_bind(&j.a, b.dev);
    pl->blobs.push_back(std::move(b));
    pl->plain_dec.push_back(PlainDecode{j.a, j.phase, j.k, j.mod, j.valid});
  } else {
    pl->dec.host.push_back(j);
    layout_decode(pl, pl->dec_rows);
  }
  if (gray) {
    double* ph2 = (double*)pl->arena.take(np * 8);
    int32_t* k2 = (int32_t*)pl->arena.take(np * 4);
    pl->gray_fix.push_back(j.a);
    pl->gray_in_ph.push_back(phase_raw);
    pl->gray_out_ph.push_back(ph2);
    pl->gray_in_k.push_back(k);
    pl->gray_out_k.push_back(k2);
    pl->gray_valid.push_back(j.valid);
    *ph_for_match = ph2;
    *k_for_match = k2;
  }
  if (hdr && dc.post_smooth) {
    double* sm = (double*)pl->arena.take(np * 8);
    pl->hdr_smooth.push_back(HdrSmooth{*ph_for_match, j.valid, sm, rows, cols});
    *ph_for_match = sm;
    *smooth_in_match = 0;
  }
  return PSGC_OK;
}

}  // namespace psgc

using namespace psgc;

struct psgc_batch {
  psgc_ctx* ctx = nullptr;
  Plan plan;
  bool graph = false;
};

// Captures the plan's launches into a graph (replacing an earlier one). The
// launches are recorded on the context's own stream (the legacy default
// stream, which the context may be bound to, cannot be captured); replays go
// to the bound stream.
static int capture(psgc_batch* b) {
  hipStream_t s = b->ctx->own_stream;
  Plan& pl = b->plan;
  if (pl.exec) hipGraphExecDestroy(pl.exec);
  if (pl.graph) hipGraphDestroy(pl.graph);
  pl.exec = nullptr;
  pl.graph = nullptr;
  b->graph = false;
  const bool prof = g_prof_on;
  g_prof_on = false;  // no event records inside the captured graph
  int rc = PSGC_OK;
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = run_plan(&pl, s);
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(s, &graph);
    if (rc == PSGC_OK && e2 == hipSuccess) {
      pl.graph = graph;
      e = hipGraphInstantiate(&pl.exec, graph, nullptr, nullptr, 0);
    } else {
      e = e2;
      if (graph) hipGraphDestroy(graph);
    }
  }
  g_prof_on = prof;
  if (rc != PSGC_OK) return rc;
  if (e != hipSuccess) PSGC_FAIL(PSGC_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
  b->graph = true;
  return PSGC_OK;
}

extern "C" {

int psgc_batch_create(psgc_ctx* ctx, const psgc_pair_desc* pairs, int n, int flags,
                      psgc_batch** out) {
  if (!ctx || !pairs || n < 1 || !out) PSGC_FAIL(PSGC_E_INVALID, "bad batch arguments");
  *out = nullptr;
  PSGC_HIP(hipSetDevice(ctx->device));
  for (int i = 0; i < n; ++i) {
    const psgc_pair_desc& p = pairs[i];
    if (p.rows < 1 || p.cols < 1 || !p.xyz || !p.P)
      PSGC_FAIL(PSGC_E_INVALID, "pair %d: bad size or null output", i);
    if (p.n_frames < 1 || p.n_frames > 64) PSGC_FAIL(PSGC_E_INVALID, "pair %d: n_frames", i);
    if (p.pha_f < 0 || p.pha_f > 65535) PSGC_FAIL(PSGC_E_INVALID, "pair %d: pha_f", i);
  }
  psgc_batch* b = new psgc_batch();
  b->ctx = ctx;
  b->plan.ctx = ctx;
  auto build = [&](Plan* pl) -> int {
    for (int i = 0; i < n; ++i) {
      const psgc_pair_desc& p = pairs[i];
      const size_t np = (size_t)p.rows * p.cols;
      const int sb = p.sample_type == PSGC_SAMPLE_U8 ? 1 : (p.sample_type == PSGC_SAMPLE_U16 ? 2 : 8);
      const int64_t stride = p.row_stride_bytes ? p.row_stride_bytes : (int64_t)p.cols * sb;
      double* lph = (double*)pl->arena.take(np * 8);
      double* rph = (double*)pl->arena.take(np * 8);
      int32_t* lk = (int32_t*)pl->arena.take(np * 4);
      int32_t* rk = (int32_t*)pl->arena.take(np * 4);
      ReconInput in{};
      int sl = 0, sr = 0;
      PSGC_TRY(plan_decode(pl, p.decode, p.left, p.n_frames, p.sample_type, p.rows, p.cols,
                           stride, lph, lk, nullptr, nullptr, &in.lph, &in.lk, &sl));
      PSGC_TRY(plan_decode(pl, p.decode, p.right, p.n_frames, p.sample_type, p.rows, p.cols,
                           stride, rph, rk, nullptr, nullptr, &in.rph, &in.rk, &sr));
      in.smooth = sl;  // both cameras share the decode configuration
      PSGC_TRY(plan_reconstruct(pl, i, p.recon, p.rows, p.cols, p.pha_f, p.Q,
                                p.coef[2] ? p.coef : nullptr, p.R1, in, p.xyz,
                                p.P, p.disp, nullptr, nullptr));
    }
    return PSGC_OK;
  };
  int rc = build_plan(&b->plan, build);
  if (rc != PSGC_OK) {
    delete b;
    return rc;
  }
  hipStream_t s = ctx->stream;
  if (flags & PSGC_BATCH_AUTOTUNE) {
    rc = run_plan(&b->plan, s);  // makes every stage's inputs valid
    if (rc == PSGC_OK) rc = autotune(&b->plan, s);
    if (rc != PSGC_OK) {
      delete b;
      return rc;
    }
  }
  if (flags & PSGC_BATCH_GRAPH) {
    rc = capture(b);
    if (rc != PSGC_OK) {
      delete b;
      return rc;
    }
  }
  *out = b;
  return PSGC_OK;
}

int psgc_batch_run(psgc_batch* b) {
  if (!b) PSGC_FAIL(PSGC_E_INVALID, "null batch");
  PSGC_HIP(hipSetDevice(b->ctx->device));
  if (b->graph) {
    PSGC_HIP(hipGraphLaunch(b->plan.exec, b->ctx->stream));
    return PSGC_OK;
  }
  return run_plan(&b->plan, b->ctx->stream);
}

int psgc_batch_run_eager(psgc_batch* b) {
  if (!b) PSGC_FAIL(PSGC_E_INVALID, "null batch");
  PSGC_HIP(hipSetDevice(b->ctx->device));
  return run_plan(&b->plan, b->ctx->stream);
}

int psgc_batch_destroy(psgc_batch* b) {
  if (!b) return PSGC_OK;
  hipSetDevice(b->ctx->device);
  hipStreamSynchronize(b->ctx->stream);
  delete b;
  return PSGC_OK;
}

// The row-match group named "match@<width>" or "match_threads@<width>".
static MatchGroup* find_group(Plan* pl, const std::string& st) {
  const size_t at = st.find('@');
  char* end = nullptr;
  const long w = std::strtol(st.c_str() + at + 1, &end, 10);
  if (end == st.c_str() + at + 1 || *end != 0) return nullptr;
  for (MatchGroup& mg : pl->match_groups)
    if (mg.cols == w) return &mg;
  return nullptr;
}

int psgc_batch_grid(psgc_batch* b, const char* stage, int* grid) {
  if (!b || !stage || !grid) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  const std::string st = stage;
  if (st == "decode_rows") {
    *grid = b->plan.dec_rows;
    return PSGC_OK;
  }
  if (st == "mls_variant") {
    *grid = b->plan.mls_variant;
    return PSGC_OK;
  }
  if (st == "match_threads") {
    *grid = b->plan.match_threads;
    return PSGC_OK;
  }
  if (st == "match_groups") {
    *grid = b->plan.match_grouped ? (int)b->plan.match_groups.size() : 0;
    return PSGC_OK;
  }
  if (st.rfind("match@", 0) == 0 || st.rfind("match_threads@", 0) == 0) {
    MatchGroup* mg = find_group(&b->plan, st);
    if (!mg) PSGC_FAIL(PSGC_E_INVALID, "no row-match group '%s'", stage);
    *grid = st[5] == '@' ? mg->grid : mg->threads;
    return PSGC_OK;
  }
  auto it = b->plan.grid.find(st);
  *grid = it != b->plan.grid.end() ? it->second : 0;
  return PSGC_OK;
}

int psgc_batch_set_grid(psgc_batch* b, const char* stage, int value) {
  if (!b || !stage) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  const std::string st = stage;
  static const char* const kStages[] = {"decode", "match", "median5", "repro_gauss",
                                        "sor_knn", "mls"};
  PSGC_HIP(hipSetDevice(b->ctx->device));
  if (st == "decode_rows") {
    if (value != 2 && value != 4)
      PSGC_FAIL(PSGC_E_INVALID, "decode_rows must be 2 or 4 (got %d)", value);
    if (!b->plan.dec.host.empty()) PSGC_TRY(relayout_decode(&b->plan, value, b->ctx->stream));
    else b->plan.dec_rows = value;
    b->plan.grid["decode_rows"] = value;
  } else if (st == "match_threads") {
    if (value != 128 && value != 256 && value != 512)
      PSGC_FAIL(PSGC_E_INVALID, "match_threads must be 128, 256 or 512 (got %d)", value);
    b->plan.match_threads = value;
    b->plan.grid["match_threads"] = value;
  } else if (st == "match_groups") {
    if (value != 0 && value != 1) PSGC_FAIL(PSGC_E_INVALID, "match_groups must be 0 or 1");
    if (value && b->plan.match_groups.empty())
      PSGC_FAIL(PSGC_E_INVALID, "the batch has rows of one width only");
    b->plan.match_grouped = value != 0;
  } else if (st.rfind("match@", 0) == 0 || st.rfind("match_threads@", 0) == 0) {
    MatchGroup* mg = find_group(&b->plan, st);
    if (!mg) PSGC_FAIL(PSGC_E_INVALID, "no row-match group '%s'", stage);
    if (st[5] == '@') {
      if (value < 1 || value > (1 << 20))
        PSGC_FAIL(PSGC_E_INVALID, "grid of %d workgroups for '%s' (1..2^20)", value, stage);
      mg->grid = value;
    } else {
      if (value != 128 && value != 256 && value != 512)
        PSGC_FAIL(PSGC_E_INVALID, "%s must be 128, 256 or 512 (got %d)", stage, value);
      mg->threads = value;
    }
  } else if (st == "mls_variant") {
    if (value != kMlsPlain && value != kMlsCached)
      PSGC_FAIL(PSGC_E_INVALID, "mls_variant must be 0 or 1 (got %d)", value);
    b->plan.mls_variant = value;
    b->plan.grid["mls_variant"] = value;
  } else {
    bool known = false;
    for (const char* k : kStages) known = known || st == k;
    if (!known) PSGC_FAIL(PSGC_E_INVALID, "unknown stage '%s'", stage);
    if (value < 1 || value > (1 << 20))
      PSGC_FAIL(PSGC_E_INVALID, "grid of %d workgroups for '%s' (1..2^20)", value, stage);
    b->plan.grid[st] = value;
  }
  if (b->graph) PSGC_TRY(capture(b));
  return PSGC_OK;
}

}  // extern "C"

namespace psgc {

// Tiled single-pair paths used by psgc_decode / psgc_reconstruct.
int tiled_decode(psgc_ctx* ctx, const DecodeArgs& a, double* phase_raw, int32_t* k, double* mod,
                 uint8_t* valid, hipStream_t s) {
  DecodeJob j{};
  j.a = a;
  j.phase = phase_raw;
  j.k = k;
  j.mod = mod;
  j.valid = valid;
  j.tiles_x = tiles(a.cols, kDecodeTileCols);
  // 2-row tiles: the batch autotuner's choice on full-size pairs
  j.tiles_y = tiles(a.rows, 2);
  j.tile_base = 0;
  void* d;
  PSGC_TRY(ctx_buf(ctx, S_QR, sizeof(DecodeJob), &d));
  PSGC_HIP(hipMemcpyAsync(d, &j, sizeof(DecodeJob), hipMemcpyHostToDevice, s));
  return launch_decode_tiled((const DecodeJob*)d, 1, (int64_t)j.tiles_x * j.tiles_y,
                             default_grid(ctx, 64), 2, s);
}

int tiled_reconstruct(psgc_ctx* ctx, const psgc_recon_config& rc, const double* lp,
                      const double* rp, const int32_t* lk, const int32_t* rk, int rows, int cols,
                      int pha_f, const double* Q, const float* const* coef, const double* R1,
                      float* xyz, uint8_t* P, float* disp, int32_t* mc, hipStream_t s) {
  Plan* pl = new Plan();
  pl->ctx = ctx;
  ReconInput in{lp, rp, lk, rk, 0};
  auto build = [&](Plan* p) -> int {
    return plan_reconstruct(p, 0, rc, rows, cols, pha_f, Q, coef, R1, in, xyz, P, disp, mc,
                            nullptr);
  };
  int r = build_plan(pl, build);
  if (r == PSGC_OK) r = run_plan(pl, s);
  // the plan's arena must outlive the queued kernels
  if (r == PSGC_OK && hipStreamSynchronize(s) != hipSuccess) r = PSGC_E_HIP;
  delete pl;
  return r;
}

}  // namespace psgc

namespace psgc {

namespace {
int upload_job(psgc_ctx* ctx, const CloudJob& j, CloudJob** out, hipStream_t s) {
  void* d;
  PSGC_TRY(ctx_buf(ctx, S_QR, sizeof(CloudJob), &d));
  PSGC_HIP(hipMemcpyAsync(d, &j, sizeof(CloudJob), hipMemcpyHostToDevice, s));
  *out = (CloudJob*)d;
  return PSGC_OK;
}

CloudJob base_job(int rows, int cols) {
  CloudJob c{};
  c.rows = rows;
  c.cols = cols;
  c.tiles_x = tiles(cols, 16);
  c.tiles_y = tiles(rows, 16);
  c.tile_base = 0;
  return c;
}
}  // namespace

int tiled_median(psgc_ctx* ctx, const float* src, float* dst, int rows, int cols, hipStream_t s) {
  CloudJob c = base_job(rows, cols);
  c.disp_in = src;
  c.disp_med = dst;
  CloudJob* d;
  PSGC_TRY(upload_job(ctx, c, &d, s));
  PSGC_TRY(launch_median_tiled(d, 1, (int64_t)c.tiles_x * c.tiles_y, default_grid(ctx, 64), s));
  return hipStreamSynchronize(s) == hipSuccess ? PSGC_OK : PSGC_E_HIP;
}

int tiled_sor(psgc_ctx* ctx, const float* xyz, const uint8_t* P, float* dst, uint8_t* dstP,
              int rows, int cols, int k, float std_mul, int half, double* md, double* partial,
              double* stats, hipStream_t s) {
  if (half > kSorTiledMaxHalf) return launch_sor(xyz, P, dst, dstP, rows, cols, k, std_mul, half, md, partial,
                                  stats, s);
  if (k < 1 || k > 64) PSGC_FAIL(PSGC_E_INVALID, "sor k=%d outside [1, 64]", k);
  CloudJob c = base_job(rows, cols);
  c.sor_xyz = xyz;
  c.sor_P = P;
  c.sor_k = k;
  c.sor_half = half;
  c.md = md;
  CloudJob* d;
  PSGC_TRY(upload_job(ctx, c, &d, s));
  PSGC_TRY(launch_sor_knn_tiled(d, 1, (int64_t)c.tiles_x * c.tiles_y, default_grid(ctx, 256), half,
                                k, k, s));
  PSGC_TRY(launch_sor_finish(xyz, md, dst, dstP, (size_t)rows * cols, std_mul, partial, stats, s));
  return hipStreamSynchronize(s) == hipSuccess ? PSGC_OK : PSGC_E_HIP;
}

int tiled_mls(psgc_ctx* ctx, const float* src, const uint8_t* P, float* dst, float* tmp, int rows,
              int cols, int hw, float sigma, int order, float r2, int iterations, int semantics,
              hipStream_t s) {
  if (hw > kMlsTiledMaxHalf) {
    const float* in = src;
    for (int it = 0; it < iterations; ++it) {
      float* out = ((iterations - 1 - it) % 2 == 0) ? dst : tmp;
      PSGC_TRY(launch_mls(in, P, out, rows, cols, hw, sigma, order, r2, semantics, s));
      in = out;
    }
    return PSGC_OK;
  }
  const float* in = src;
  for (int it = 0; it < iterations; ++it) {
    float* out = ((iterations - 1 - it) % 2 == 0) ? dst : tmp;
    CloudJob c = base_job(rows, cols);
    c.mls_src = in;
    c.mls_P = P;
    c.mls_dst = out;
    c.mls_hw = hw;
    c.mls_order = order;
    c.mls_sem = semantics;
    c.mls_inv2s2 = 1.0f / (2.0f * sigma * sigma);
    c.mls_r2 = r2;
    CloudJob* d;
    PSGC_TRY(upload_job(ctx, c, &d, s));
    PSGC_TRY(launch_mls_tiled(d, 1, (int64_t)c.tiles_x * c.tiles_y, default_grid(ctx, 128),
                              mls_class(hw), kMlsPlain, mls_sem_bit(semantics), s));
    // the job buffer is reused by the next iteration
    PSGC_HIP(hipStreamSynchronize(s));
    in = out;
  }
  return PSGC_OK;
}

}  // namespace psgc
