/*
 * psgc.h — C ABI of the MI355X (gfx950) reconstruction path for binocular
 * phase-shift structured light: wrapped-phase decoding, three-frequency
 * heterodyne or Gray-code unwrapping, order-constrained row matching with a
 * sub-pixel root, Q-matrix triangulation and the organized point-cloud filters
 * (Gaussian, radius outlier, statistical outlier, MLS).
 *
 * Reference seams this header replaces (paths relative to
 * PhaseShiftGrayCode3D/src/ of the reference):
 *   JSLAB_DecodePhase_GPU / _Warmup / _Release
 *       JSLAB_PhaseShiftDecoder_V2_GPU.h:50-60
 *   JSLAB_ReconstructPipeline_GPU
 *       JSLAB_PhaseEpipolarReconstructor_V2_GPU.h:131-137
 *   applyMLSFilterToPointCloud_GPU / applyMLSFilterToPointCloud
 *       JSLAB_MLS_Filter.h:42-57
 *   the CPU-side building blocks that the GPU seams do not cover
 *   (JSLAB_PhaseEpipolarReconstructor_V2.cpp:79-195, 397-480, 933-1006)
 *
 * Conventions
 *   - extern "C", plain pointers and sizes; no exceptions cross the ABI.
 *   - Every entry point returns an int status: PSGC_OK (0) or a negative
 *     PSGC_E_* code; psgc_last_error() returns a thread-local message.
 *   - Image-shaped buffers are dense row-major (row stride = cols elements),
 *     except the input frames, which take an explicit row stride in bytes.
 *   - `mem` selects where the caller's buffers live: PSGC_MEM_DEVICE (HBM,
 *     the fast path; the call is asynchronous on the context stream) or
 *     PSGC_MEM_HOST (the library stages through context-owned device buffers
 *     and the call returns after the results are back in host memory, as the
 *     reference launchers do).
 *   - One context per thread; a context owns its stream, staging buffers and
 *     captured graphs (the reference keeps process-global static buffers,
 *     JSLAB_PhaseShiftDecoder_V2_GPU.cu:153-158, so it is not re-entrant).
 */
#ifndef PSGC_H
#define PSGC_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define PSGC_API __attribute__((visibility("default")))
#else
#define PSGC_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define PSGC_ABI_VERSION 1

/* ---- status codes -------------------------------------------------------- */
enum {
    PSGC_OK = 0,
    PSGC_E_INVALID = -1,   /* null pointer, bad size, bad configuration       */
    PSGC_E_NODEVICE = -2,  /* no HIP device, or device index out of range      */
    PSGC_E_ALLOC = -3,     /* device or host allocation failed                 */
    PSGC_E_HIP = -4,       /* a HIP runtime call failed                        */
    PSGC_E_UNSUPPORTED = -5
};

enum { PSGC_MEM_DEVICE = 0, PSGC_MEM_HOST = 1 };

/* Sample type of the input frames. The reference CPU decoder accepts CV_8U
 * and CV_64F (JSLAB_PhaseShiftDecoder_V2.cpp:260-261); 12-bit cameras deliver
 * uint16, which is read natively here (each sample converts exactly to f64). */
enum { PSGC_SAMPLE_U8 = 0, PSGC_SAMPLE_U16 = 1, PSGC_SAMPLE_F64 = 2 };

/* Unwrapping schedule. */
enum {
    PSGC_UNWRAP_HETERODYNE3 = 0, /* three-frequency heterodyne (reference V2)  */
    PSGC_UNWRAP_GRAYCODE = 1     /* one phase-shift frequency + Gray code with
                                    complement frames (textbook; not in the
                                    reference)                                 */
};

/* Which of the reference's two behaviours to follow where its CPU and GPU
 * paths disagree (JSLAB_Dual_GPU移植说明.md:68-78):
 *   CPU: point-cloud Gaussian keeps the mask and divides by multiplying with
 *        the reciprocal (JSLAB_PhaseEpipolarReconstructor_V2.cpp:79-124); the
 *        radius outlier filter reads the pre-Gaussian mask (:466); MLS falls
 *        back straight to order 0 (JSLAB_MLS_Filter.cpp:277-279).
 *   GPU: Gaussian clears the mask of rejected points and divides
 *        (JSLAB_PhaseEpipolarReconstructor_V2_GPU.cu:220-274); the outlier
 *        filter reads that mask (:476-485); MLS falls back 2 -> 1 -> 0
 *        (JSLAB_MLS_Filter.cu:283-286). */
enum { PSGC_SEM_CPU = 0, PSGC_SEM_GPU = 1 };

/* Outlier stage after the Gaussian. */
enum { PSGC_OUTLIER_NONE = 0, PSGC_OUTLIER_ROR = 1, PSGC_OUTLIER_SOR = 2 };

/* ---- decoder configuration ----------------------------------------------
 * Fields 1..9 mirror JSLAB_PhaseShiftDecoder::Config
 * (JSLAB_PhaseShiftDecoder.h:22-54) with the same meaning and defaults. */
typedef struct psgc_decode_config {
    double freq[3];               /* fringe counts of the three levels        */
    int32_t shift_num[3];         /* phase steps per level (3..12)            */
    double modulation_threshold;  /* default 1.0                              */
    double preblur_sigma;         /* default 3.0; <=0: fixed [1/4,1/2,1/4]    */
    int32_t avg_three_phase;      /* default 1                                */
    int32_t post_smooth;          /* default 1                                */
    int32_t enable_hdr;           /* default 0                                */
    double hdr_saturation_gray;   /* default 255                              */
    double hdr_min_amplitude;     /* default 2                                */
    /* additions (no reference counterpart) */
    int32_t unwrap_mode;          /* PSGC_UNWRAP_*; default heterodyne         */
    int32_t gray_bits;            /* Gray-code bits (graycode mode), e.g. 6    */
    double gray_period_px;        /* projector pixels per phase period          */
    double gray_min_contrast;     /* |I - I_complement| below => invalid        */
} psgc_decode_config;

PSGC_API void psgc_decode_config_default(psgc_decode_config* cfg);

/* ---- reconstructor configuration -----------------------------------------
 * Fields up to mls_iterations mirror JSLAB_PhaseEpipolarReconstructor_V2::Config
 * (JSLAB_PhaseEpipolarReconstructor_V2.h:18-55); the reference's GPU class
 * has the same fields with other defaults
 * (JSLAB_PhaseEpipolarReconstructor_V2_GPU.h:27-61). */
typedef struct psgc_recon_config {
    double min_depth;             /* 0                                        */
    double max_depth;             /* 1000                                     */
    double min_phase_diff;        /* 0.03 (unused by the reference's path)     */
    double max_phase_diff;        /* 0.4  (unused by the reference's path)     */
    double grad_threshold;        /* 1.0: max |phiR - phiL| of a match         */
    int32_t use_rectified_coords; /* 1                                        */
    int32_t use_disp_median_filter; /* 1                                      */
    int32_t use_full_left_rectification; /* 1 (host-side setup only)          */
    float radius_search;          /* CPU 0.7 / GPU 0.05                        */
    int32_t radius_min_neighbors; /* CPU 10  / GPU 5                           */
    int32_t radius_window_size;   /* CPU 10  / GPU 5 (half = size / 2)         */
    int32_t use_mls;              /* CPU 0   / GPU 1                           */
    int32_t mls_window_size;      /* half window: CPU 30 / GPU 25              */
    float mls_spatial_sigma;      /* CPU 0.6 / GPU 0.5                         */
    int32_t use_gpu_mls;          /* 1                                        */
    int32_t mls_polynomial_order; /* 1                                        */
    float mls_max_radius;         /* 0 = 3 sigma                               */
    int32_t mls_iterations;       /* CPU 3   / GPU 2                           */
    /* additions */
    int32_t semantics;            /* PSGC_SEM_CPU / PSGC_SEM_GPU               */
    int32_t gauss_window;         /* 3 in the reference                        */
    float gauss_sigma;            /* 0.7 in the reference                      */
    int32_t outlier_mode;         /* PSGC_OUTLIER_*; reference = ROR           */
    int32_t sor_k;                /* neighbours averaged by SOR                */
    float sor_std_mul;            /* threshold = mean + std_mul * stddev       */
    int32_t sor_window;           /* half window of the organized kNN search   */
} psgc_recon_config;

/* Defaults of the reference CPU class (semantics = PSGC_SEM_CPU). */
PSGC_API void psgc_recon_config_default(psgc_recon_config* cfg);
/* Defaults of the reference GPU class (semantics = PSGC_SEM_GPU). */
PSGC_API void psgc_recon_config_default_gpu(psgc_recon_config* cfg);

/* ---- context ------------------------------------------------------------- */
typedef struct psgc_ctx psgc_ctx;

PSGC_API const char* psgc_last_error(void);
PSGC_API int psgc_abi_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
PSGC_API int psgc_device_count(void);

/* Creates a context on `device` with its own stream. */
PSGC_API int psgc_create(psgc_ctx** out, int device);
PSGC_API int psgc_destroy(psgc_ctx* ctx);
/* Use the caller's hipStream_t (e.g. torch's current stream) instead of the
 * context's own; NULL restores the context's stream. The legacy default
 * stream is given as hipStreamLegacy ((hipStream_t)1): work on it is ordered
 * with every blocking stream's, and so with other contexts bound to it. */
PSGC_API int psgc_set_stream(psgc_ctx* ctx, void* hip_stream);
PSGC_API int psgc_synchronize(psgc_ctx* ctx);

/* Replaces JSLAB_DecodePhase_GPU_Warmup (JSLAB_PhaseShiftDecoder_V2_GPU.h:56):
 * initialises the HIP runtime on the current device. */
PSGC_API int psgc_warmup(void);
/* Replaces JSLAB_DecodePhase_GPU_Release (JSLAB_PhaseShiftDecoder_V2_GPU.h:60):
 * frees the context's cached staging buffers and graphs. */
PSGC_API int psgc_release(psgc_ctx* ctx);

/* ---- decode --------------------------------------------------------------
 * Replaces JSLAB_DecodePhase_GPU (JSLAB_PhaseShiftDecoder_V2_GPU.h:50-52) and
 * JSLAB_PhaseShiftDecoderV2::setImages + startCalculate + getOrderMap +
 * getModulationMap (JSLAB_PhaseShiftDecoder_V2.cpp:256-310).
 *   frames[n_frames]: one plane per frame, rows x cols samples, row stride
 *     `row_stride_bytes` (0 = dense). Heterodyne: frames ordered level 0
 *     steps, level 1 steps, level 2 steps (sum of shift_num). Gray code:
 *     shift_num[0] phase frames, then gray_bits pattern frames, then
 *     gray_bits complement frames (MSB first).
 *   phase (f64, post-smoothed if post_smooth), k (i32 fringe order),
 *   mod (f64 normalized modulation of level 0; may be NULL),
 *   valid (u8, 255 valid / 0 invalid; may be NULL).
 * Unlike the reference GPU launcher (which saturates everything to 8 bit,
 * JSLAB_PhaseShiftDecoder_V2_GPU.cu:252-258) every sample type is decoded
 * exactly as the reference CPU decoder would after an exact f64 conversion. */
PSGC_API int psgc_decode(psgc_ctx* ctx, const psgc_decode_config* cfg,
                const void* const* frames, int n_frames, int sample_type,
                int rows, int cols, int64_t row_stride_bytes,
                double* phase, int32_t* k, double* mod, uint8_t* valid,
                int mem);

/* ---- reconstruct ---------------------------------------------------------
 * Replaces JSLAB_ReconstructPipeline_GPU
 * (JSLAB_PhaseEpipolarReconstructor_V2_GPU.h:131-137) and
 * JSLAB_PhaseEpipolarReconstructor_V2::startCalculate
 * (JSLAB_PhaseEpipolarReconstructor_V2.cpp:397-480): order-constrained match
 * -> optional 5x5 median -> Q reprojection -> Gaussian -> outlier filter
 * -> optional MLS. Inputs are rectified maps (remap is the caller's, as in
 * the reference, where it stays on the host).
 *   pha_f = number of fringe orders (the reference passes freq[0]).
 *   Q: 4x4 row-major, R1: 3x3 row-major (f64).
 *   xyz: rows*cols*3 f32 (camera or rectified coordinates), P: u8 0/1,
 *   disp (may be NULL): f32 disparity after match (+median), 0 = none,
 *   match_col (may be NULL): i32 right column of the integer match, -1 none. */
PSGC_API int psgc_reconstruct(psgc_ctx* ctx, const psgc_recon_config* cfg,
                     const double* l_phase, const double* r_phase,
                     const int32_t* l_k, const int32_t* r_k,
                     int rows, int cols, int pha_f,
                     const double Q[16], const double R1[9],
                     float* xyz, uint8_t* P, float* disp, int32_t* match_col,
                     int mem);

/* ---- individual stages (stage-isolated parity, as the reference's compare
 * harness does, JSLAB_Demo_Dual.cpp:448-513) ------------------------------ */

/* stereoMatch (JSLAB_PhaseEpipolarReconstructor_V2.cpp:933-1006): disp is
 * f64 (0 = no match); match_col may be NULL. */
PSGC_API int psgc_stereo_match(psgc_ctx* ctx, const double* l_phase,
                      const double* r_phase, const int32_t* l_k,
                      const int32_t* r_k, int rows, int cols, int pha_f,
                      double grad_threshold, double* disp, int32_t* match_col,
                      int mem);

/* 5x5 median with replicated border on f32 (cv::medianBlur(…, 5),
 * JSLAB_PhaseEpipolarReconstructor_V2.cpp:405-412). */
PSGC_API int psgc_median5(psgc_ctx* ctx, const float* src, float* dst, int rows,
                 int cols, int mem);

/* Q reprojection (JSLAB_PhaseEpipolarReconstructor_V2.cpp:414-461); disp is
 * f32 as the reference stores it after the median. */
PSGC_API int psgc_reproject(psgc_ctx* ctx, const float* disp, int rows, int cols,
                   const double Q[16], const double R1[9], double min_depth,
                   double max_depth, int use_rectified_coords, float* xyz,
                   uint8_t* P, int mem);

/* Per-pixel polynomial triangulation: evaluateV2CalibPoint and the point
 * loop of startCalculate_v2 (JSLAB_PhaseEpipolarReconstructor_V2.cpp:13-76,
 * 493-545). coef_x/y/z: the _rx/_ry/_rz calibration maps
 * (JSLAB_StereoCalibParams_V2.cpp:36-41), rows*cols*4 f32 each (CV_32FC4):
 * markers -1..-4 in the fourth term of x or z select the form of
 * (x, y, z)(d). disp f64; R1 rotates to camera coordinates when
 * use_rectified_coords is 0; the depth gate is the Q path's. */
PSGC_API int psgc_reproject_poly(psgc_ctx* ctx, const double* disp, int rows, int cols,
                                 const float* coef_x, const float* coef_y, const float* coef_z,
                                 const double R1[9], double min_depth, double max_depth,
                                 int use_rectified_coords, float* xyz, uint8_t* P, int mem);

/* startCalculate_v2 (JSLAB_PhaseEpipolarReconstructor_V2.cpp:482-593):
 * psgc_reconstruct with the per-pixel polynomial triangulation in place of
 * the Q reprojection (same configuration, outputs and staging rules). */
PSGC_API int psgc_reconstruct_poly(psgc_ctx* ctx, const psgc_recon_config* cfg,
                                   const double* l_phase, const double* r_phase,
                                   const int32_t* l_k, const int32_t* r_k, int rows, int cols,
                                   int pha_f, const float* coef_x, const float* coef_y,
                                   const float* coef_z, const double R1[9], float* xyz,
                                   uint8_t* P, float* disp, int32_t* match_col, int mem);

/* ROI-restricted reconstruction: startCalculate_v2_roi(xyz, P, RectifiedROI)
 * (JSLAB_PhaseEpipolarReconstructor_V2.cpp:693-914). left_rect and
 * right_rect = {x, y, width, height} in rectified pixels with the same y and
 * height (RectifiedROI, …V2.h:97-102); mask = width x height u8 of
 * left_rect (0: outside the ROI). The full-size maps are read in place.
 * The left crop is matched against the right crop under the mask (edge test
 * in right-crop columns, disparity from global columns), then median,
 * polynomial triangulation with the coefficients at global pixels,
 * Gaussian and the outlier stage; with ROR, a result that keeps fewer than
 * 10 % of the points is replaced by the Gaussian points and the pre-ROR
 * mask (…V2.cpp:889-901); then MLS. xyz, P and disp (f32, after the
 * median; may be NULL) have left_rect's size. */
PSGC_API int psgc_reconstruct_poly_roi(psgc_ctx* ctx, const psgc_recon_config* cfg,
                                       const double* l_phase, const double* r_phase,
                                       const int32_t* l_k, const int32_t* r_k, int rows,
                                       int cols, int pha_f, const int32_t left_rect[4],
                                       const int32_t right_rect[4], const uint8_t* mask,
                                       const float* coef_x, const float* coef_y,
                                       const float* coef_z, const double R1[9], float* xyz,
                                       uint8_t* P, float* disp, int mem);

/* The right search window of prepareRectifiedROI (…V2.cpp:629-669): the
 * fringe orders 1..pha_f present under the mask in left_rect, then the first
 * and last column of the right order map (left_rect's rows) carrying one of
 * them, padded by 32 px and clipped to the image; with none, the full rows.
 * right_rect (host) receives {x, y, width, height}. Synchronous. */
PSGC_API int psgc_roi_right_window(psgc_ctx* ctx, const int32_t* l_k, const int32_t* r_k,
                                   int rows, int cols, int pha_f, const int32_t left_rect[4],
                                   const uint8_t* mask, int32_t right_rect[4], int mem);

/* Reads one coefficient map file (JSLAB_StereoCalibParams_V2::loadMatBinary,
 * JSLAB_StereoCalibParams_V2.cpp:4-17): int32 type, rows, cols, then the
 * row-major data; the type must be CV_32FC4 (29). With out == NULL only
 * *rows and *cols are set; otherwise out receives rows*cols*4 f32 (host). */
PSGC_API int psgc_load_coef_bin(const char* path, int* rows, int* cols, float* out);

/* Point-cloud Gaussian (JSLAB_PhaseEpipolarReconstructor_V2.cpp:79-124):
 * window `win` (odd), sigma; with PSGC_SEM_GPU dst_P receives the updated
 * mask, with PSGC_SEM_CPU dst_P receives a copy of P. */
PSGC_API int psgc_filter_gaussian(psgc_ctx* ctx, const float* xyz, const uint8_t* P,
                         float* dst_xyz, uint8_t* dst_P, int rows, int cols,
                         int win, float sigma, int semantics, int mem);

/* Radius outlier removal (JSLAB_PhaseEpipolarReconstructor_V2.cpp:126-195). */
PSGC_API int psgc_filter_ror(psgc_ctx* ctx, const float* xyz, const uint8_t* P,
                    float* dst_xyz, uint8_t* dst_P, int rows, int cols,
                    float radius, int min_neighbors, int window_size, int mem);

/* Statistical outlier removal (textbook/PCL definition, no reference code):
 * for every valid point, mean Euclidean distance to its k nearest valid
 * neighbours inside a (2*half_window+1)^2 pixel window; points with fewer
 * than k candidates are removed and excluded from the statistics; keep
 * if mean_dist <= mean + std_mul * stddev (sample stddev over kept-eligible
 * points). mean_dist (may be NULL) receives the per-point distance (f64,
 * 0 where not computed); stats (may be NULL) receives {count, mean, stddev,
 * threshold}. */
PSGC_API int psgc_filter_sor(psgc_ctx* ctx, const float* xyz, const uint8_t* P,
                    float* dst_xyz, uint8_t* dst_P, int rows, int cols, int k,
                    float std_mul, int half_window, double* mean_dist,
                    double stats[4], int mem);

/* MLS projection (applyMLSFilterToPointCloud_GPU, JSLAB_MLS_Filter.h:42-47;
 * semantics selects the order-fallback rule). */
PSGC_API int psgc_filter_mls(psgc_ctx* ctx, const float* src_xyz, const uint8_t* P,
                    float* dst_xyz, int rows, int cols, int window_size,
                    float spatial_sigma, int polynomial_order,
                    float max_radius, int iterations, int semantics, int mem);

/* ---- unorganized clouds (a merged multi-view cloud) -----------------------
 * No counterpart in the reference, which filters organized clouds only; the
 * per-point rules are its organized filters' with the window replaced by a
 * neighbour search. xyz: n points (f32 x, y, z interleaved); valid: one u8
 * per point or NULL (all). A point takes part when valid and finite; the
 * others pass through (MLS, Gaussian) or are removed (SOR).
 * Neighbours come from a voxel grid of edge `voxel` (points sorted by cell;
 * a workgroup stages the (2R+1)^3 cells around one cell in LDS); sums run
 * over the cells dz, dy, dx = -R..R (z outermost), each cell's points in
 * index order, so results are deterministic. SOR: R = 1 first. MLS and
 * Gaussian: R = 1 when voxel >= the radius, R = 2 when voxel >= half of it
 * (fewer candidates: 15.6 r^3 of space instead of 27 r^3), else an error.
 *
 * psgc_cloud_sor: mean distance to the k (1..64) nearest other points,
 *   exact: the search widens ring by ring (R = 1..rmax cells, rmax <= 16)
 *   until the k-th squared distance is at most 0.99999 (R voxel)^2; a point
 *   not settled by rmax is not eligible. keep[i] = eligible and mean_dist <=
 *   mean + std_mul * sd over the eligible points (sample sd; the reduction
 *   of psgc_filter_sor). mean_dist (required): f64 per point, 0 where not
 *   eligible; stats: {count, mean, sd, threshold}. keep may be NULL.
 * psgc_cloud_mls: the MLS fit of psgc_filter_mls (orders, fallback,
 *   semantics) over the points within max_radius (3 sigma if 0), the point
 *   itself included; voxel >= half that radius. Each iteration rebuilds the
 *   grid from the moved points.
 * psgc_cloud_gaussian: weighted mean of the points within radius (itself
 *   included), weights exp(-d^2 / (2 sigma^2)), FP32; voxel >= radius / 2. */
PSGC_API int psgc_cloud_sor(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n,
                            int k, float std_mul, float voxel, int rmax, uint8_t* keep,
                            double* mean_dist, double* stats, int mem);
/* The autotuner of the k-nearest search: times psgc_cloud_sor (grid build
 * included) on this cloud at each candidate voxel (HIP events on the
 * context's stream) and returns the fastest in *best_voxel; times_ms (may be
 * NULL) receives each candidate's time. Device pointers. Every candidate
 * gives the same result for points whose k-th nearest neighbour lies within
 * rmax times the smallest candidate (the eligibility limit scales with the
 * voxel). */
PSGC_API int psgc_cloud_tune_voxel(psgc_ctx* ctx, const float* xyz, const uint8_t* valid,
                                   int64_t n, int k, int rmax, const float* candidates,
                                   int n_candidates, float* best_voxel, float* times_ms);
PSGC_API int psgc_cloud_mls(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n,
                            float spatial_sigma, int polynomial_order, float max_radius,
                            int iterations, int semantics, float voxel, float* out, int mem);
PSGC_API int psgc_cloud_gaussian(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n,
                                 float sigma, float radius, float voxel, float* out, int mem);

/* Closed-form rectification for an ideal, already rectified pair (zero
 * distortion, R = I, translation along x): fills Q (4x4) and R1 (3x3) as
 * cv::stereoRectify would with flags 0. */
PSGC_API int psgc_rectify_ideal(double fx, double fy, double cx1, double cy1,
                       double cx2, double baseline, double Q[16],
                       double R1[9]);

/* Stereo rectification for zero distortion: cv::stereoRectify as
 * setCalibrationParams calls it (JSLAB_PhaseEpipolarReconstructor_V2.cpp:
 * 302-307: flags 0, alpha -1, new size = image size), restated from OpenCV
 * 4.7.0's cvStereoRectify: each camera turns by half the relative rotation
 * (Rodrigues vector of R times -1/2), then both by the rotation that takes the
 * baseline onto the x axis (or y, when |T_y| > |T_x|); R1 = wR r^T, R2 = wR r.
 * Focal length of both rectified cameras: the mean of the two cameras' f_y
 * (f_x for vertical rigs); principal points put the mean of the four
 * projected image corners at the image centre, with the two cameras' y (x)
 * averaged. P1, P2 are 3x4, Q 4x4, row-major; P2[0][3] (P2[1][3]) =
 * t fc. K1, K2, R 3x3 row-major; T 3. Host-only; distortion must be zero
 * (the caller passes none). */
PSGC_API int psgc_stereo_rectify(const double K1[9], const double K2[9], const double R[9],
                                 const double T[3], int width, int height, double R1[9],
                                 double R2[9], double P1[12], double P2[12], double Q[16]);

/* Full-left rectification canvas: computeFullLeftRectification
 * (JSLAB_PhaseEpipolarReconstructor_V2.cpp:248-289) for zero distortion, the
 * branch setCalibrationParams takes when use_full_left_rectification is set
 * (…V2.cpp:314-329). The four corners of the width x height left image,
 * (0,0), (w-1,0), (w-1,h-1), (0,h-1) as float, go through
 * cv::undistortPoints(K1, 0, R1, P1) (normalise by K1, apply P1[:3,:3] R1,
 * divide by the third coordinate, round to float); with their float bounding
 * box [min_x, max_x] x [min_y, max_y], shift = -floor(min) is added to the
 * principal points of P1 and P2 (3x4 row-major, in place) and subtracted
 * from Q[0][3] and Q[1][3] (in place); *out_width = ceil(max_x - min_x + 1)
 * and *out_height likewise (float arithmetic), the size of the rectified
 * maps. Host-only. */
PSGC_API int psgc_full_left_rectification(const double K1[9], int width, int height,
                                          const double R1[9], double P1[12], double P2[12],
                                          double Q[16], int* out_width, int* out_height);

/* ---- kernel selection -----------------------------------------------------
 * PSGC_KERNELS_TILED (default): LDS-tiled kernels. PSGC_KERNELS_PLAIN: one
 * thread per pixel in 16x16 workgroups reading HBM directly, one launch per
 * stage, which is the structure of the reference GPU path and serves as the
 * per-pair baseline. Both produce the same integer outputs. */
enum { PSGC_KERNELS_TILED = 0, PSGC_KERNELS_PLAIN = 1 };
PSGC_API int psgc_set_kernels(psgc_ctx* ctx, int which);

/* ---- batched pipeline ------------------------------------------------------
 * Decode + reconstruct of many stereo pairs of uneven size, step count and
 * unwrapping schedule, one launch per stage for the whole batch (persistent,
 * XCD-aware grids over the concatenated tiles of all pairs), optionally
 * captured once into a hipGraph that psgc_batch_run replays. No counterpart
 * in the reference, which processes one pair per call
 * (JSLAB_DecodePhase_GPU + JSLAB_ReconstructPipeline_GPU). All pointers in a
 * pair descriptor are device pointers and must stay valid while the batch
 * exists; outputs are overwritten by every run. */
typedef struct psgc_pair_desc {
    const void* left[64];          /* frames of the left camera            */
    const void* right[64];         /* frames of the right camera           */
    int32_t n_frames;              /* per camera                           */
    int32_t sample_type;           /* PSGC_SAMPLE_*                        */
    int64_t row_stride_bytes;      /* 0 = dense                            */
    int32_t rows, cols;
    psgc_decode_config decode;
    psgc_recon_config recon;
    int32_t pha_f;                 /* fringe orders (reference: freq[0])   */
    double Q[16];
    double R1[9];
    float* xyz;                    /* rows*cols*3 f32                      */
    uint8_t* P;                    /* rows*cols u8                         */
    float* disp;                   /* optional f32 disparity (after median) */
    const float* coef[3];          /* optional: x/y/z coefficient maps (see
                                      psgc_reconstruct_poly); all NULL = Q  */
} psgc_pair_desc;

enum {
    PSGC_BATCH_GRAPH = 1,          /* capture the launches into a hipGraph  */
    PSGC_BATCH_AUTOTUNE = 2        /* time launch shapes per resolution     */
};

typedef struct psgc_batch psgc_batch;
PSGC_API int psgc_batch_create(psgc_ctx* ctx, const psgc_pair_desc* pairs, int n_pairs,
                               int flags, psgc_batch** out);
PSGC_API int psgc_batch_run(psgc_batch* batch);
PSGC_API int psgc_batch_destroy(psgc_batch* batch);
/* Runs the batch's launches directly on the stream (no graph), so that
 * per-kernel instrumentation (psgc_profile_*) can see each launch. */
PSGC_API int psgc_batch_run_eager(psgc_batch* batch);
/* Launch shape chosen for stage `stage` ("decode", "match", "median5",
 * "repro_gauss", "sor_knn", "mls"): workgroups of the persistent grid;
 * "decode_rows": rows per decode tile (2 or 4); "match_threads": threads per
 * rectified row in the row match (128, 256 or 512); "mls_variant": 0 = weights
 * evaluated in each MLS pass, 1 = the first weights kept in LDS between the
 * passes. 0 = the stage's default grid. A batch whose pairs have several
 * image widths also tunes the row match per width: "match_groups" is the
 * number of per-width launches that replace the single one (0: the single
 * launch is faster and runs), "match@<width>" and "match_threads@<width>"
 * the grid and threads per row of that width's launch. */
PSGC_API int psgc_batch_grid(psgc_batch* batch, const char* stage, int* grid);
/* Pins a launch shape (same stage names; grids 1..2^20 workgroups, decode_rows
 * 2 or 4, match_threads 128, 256 or 512, mls_variant 0 or 1, match_groups 0
 * (single launch) or 1 (per-width launches)), e.g. one saved from an earlier
 * autotuned batch of the same resolutions. A captured graph is re-captured. Results do not depend on the
 * launch shape. No counterpart in the reference (its launches have fixed
 * shapes). */
PSGC_API int psgc_batch_set_grid(psgc_batch* batch, const char* stage, int value);

/* ---- rectification, compaction and output ---------------------------------
 * psgc_rectify_maps: cv::initUndistortRectifyMap for zero distortion
 *   (JSLAB_PhaseEpipolarReconstructor_V2.cpp:333-338): for each rectified
 *   pixel, raw = K_raw R^T K_rect^-1 (u, v, 1); f32 maps rows x cols.
 * psgc_remap: cv::remap of the reference's setData (…V2.cpp:368-395):
 *   kind 0 = f64 map, INTER_LINEAR; kind 1 = i32 map, INTER_NEAREST;
 *   BORDER_CONSTANT 0. The rule is OpenCV 4.7.0's for CV_32FC1 maps: the
 *   coordinate becomes cvRound(map * 32) (half to even) for INTER_LINEAR,
 *   i.e. 1/32 px, whose integer part and fraction pick the four taps and
 *   the float weights ((1-fy)(1-fx), (1-fy)fx, fy(1-fx), fy fx), summed in
 *   that order in double; INTER_NEAREST takes cvRound(map); a tap outside
 *   the source reads 0. maps are rows x cols f32 (source pixel coordinates).
 * psgc_compact: valid points (P != 0) packed row-major into out_xyz (n*3
 *   f32) and out_gray (n u8, 255 when gray is NULL; may be NULL); *count
 *   (host) receives n. Order = the reference PLY writers' (main.cpp:85-162).
 * psgc_write_ply: binary little-endian PLY, float x y z and, when gray is
 *   given, uchar red green blue (main.cpp:132-162). Host arrays.
 * psgc_write_raw_f64: a rows x cols f64 map (phase, disparity) as raw
 *   row-major doubles, and, when info_path is not NULL, a text file naming
 *   rows, cols, type and layout (saveRaw64F, JSLAB_PhaseDisparityExporter.cpp:
 *   115-141). Host array. */
PSGC_API int psgc_rectify_maps(psgc_ctx* ctx, const double K_raw[9], const double R[9],
                               const double K_rect[9], int rows, int cols, float* map_x,
                               float* map_y, int mem);
PSGC_API int psgc_remap(psgc_ctx* ctx, int kind, const void* src, int src_rows, int src_cols,
                        const float* map_x, const float* map_y, int rows, int cols, void* dst,
                        int mem);
PSGC_API int psgc_compact(psgc_ctx* ctx, const float* xyz, const uint8_t* P,
                          const uint8_t* gray, int rows, int cols, float* out_xyz,
                          uint8_t* out_gray, int64_t* count, int mem);

/* psgc_compact_noalloc: psgc_compact without any allocation. It has no
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc). Every pointer is device memory; the call is
 * asynchronous on the context's stream, and *count (a device int64)
 * receives n. At most PSGC_COMPACT_NOALLOC_MAX_POINTS pixels per call
 * (larger calls fail with PSGC_E_INVALID). The block counts live in a static
 * per-GPU device array of 1 MiB, held until the process ends and shared by
 * every call on that GPU: calls that overlap on one GPU (other streams or
 * contexts) race and must be serialized by the caller. */
#define PSGC_COMPACT_NOALLOC_MAX_POINTS ((size_t)1 << 27)
PSGC_API int psgc_compact_noalloc(psgc_ctx* ctx, const float* xyz, const uint8_t* P,
                                  const uint8_t* gray, int rows, int cols, float* out_xyz,
                                  uint8_t* out_gray, int64_t* count);
PSGC_API int psgc_write_ply(const char* path, const float* xyz, const uint8_t* gray, int64_t n);
PSGC_API int psgc_write_raw_f64(const char* raw_path, const char* info_path, const double* data,
                                int rows, int cols);

/* ---- differentiable triangulation ----------------------------------------
 * Rectification of raw phase maps under a calibration theta (Catmull-Rom
 * bicubic resampling, so the result is C1 in theta), the sub-pixel
 * root (JSLAB_PhaseEpipolarReconstructor_V2.cpp:992-1003) with the integer
 * match held fixed, and the Q triangulation (:430-438), in FP64, with its
 * hand-written adjoint. No counterpart in the reference (it has no
 * gradients); used for calibration refinement.
 *   theta[9]  = f, cx1, cy1, cx2, cy2, baseline, wx, wy, wz (rad; right
 *               rectification rotation Rz(wz) Ry(wy) Rx(wx))
 *   nominal[5]= f0, c0Lx, c0Ly, c0Rx, c0Ry: intrinsics of the raw maps
 *   match_col = integer right column per left rectified pixel (-1: none)
 *   xyz (rows*cols*3 f64, rectified left frame), valid (u8, may be NULL):
 *               1 = triangulated, every phase sample's 4 x 4 Catmull-Rom
 *               taps within pi of their neighbours (Itoh's condition: the
 *               sample lies in smoothly unwrapped phase); 2 = triangulated,
 *               but a sample
 *               spans a phase jump of pi or more (fringe-order error, surface
 *               edge, region boundary), where xyz is not a smooth function of
 *               theta; 0 = no point (no match, taps outside the map, or a
 *               degenerate root)
 * The backward pass takes dL/dxyz and returns dL/dtheta (9 f64) and, when
 * the pointers are non-NULL, dL/dphi for both raw maps (f64, overwritten).
 * theta and nominal are host arrays; grad_theta lives where `mem` says. */
PSGC_API int psgc_adjoint_forward(psgc_ctx* ctx, const double theta[9], const double nominal[5],
                                  const double* phL, const double* phR,
                                  const int32_t* match_col, int rows, int cols, double* xyz,
                                  uint8_t* valid, int mem);
PSGC_API int psgc_adjoint_backward(psgc_ctx* ctx, const double theta[9],
                                   const double nominal[5], const double* phL,
                                   const double* phR, const int32_t* match_col, int rows,
                                   int cols, const double* grad_xyz, double* grad_theta,
                                   double* grad_phL, double* grad_phR, int mem);

/* Many views (stereo pairs) in one launch each way, e.g. per-view calibration
 * refinement of a multi-view scan. Per view, the arguments of
 * psgc_adjoint_forward / psgc_adjoint_backward; device pointers only (on the
 * context's stream; the call returns when the launch is done). grad_theta:
 * 9 doubles per view. No counterpart in the reference. */
typedef struct psgc_adjoint_view {
    double theta[9];
    double nominal[5];
    const double* phL;
    const double* phR;
    const int32_t* match_col;
    int32_t rows, cols;
    double* xyz;              /* forward: rows*cols*3                     */
    uint8_t* valid;           /* forward, optional                        */
    const double* grad_xyz;   /* backward input                           */
    double* grad_theta;       /* backward: 9                              */
    double* grad_phL;         /* backward, optional                       */
    double* grad_phR;         /* backward, optional                       */
} psgc_adjoint_view;
PSGC_API int psgc_adjoint_forward_batch(psgc_ctx* ctx, const psgc_adjoint_view* views, int n);
PSGC_API int psgc_adjoint_backward_batch(psgc_ctx* ctx, const psgc_adjoint_view* views, int n);

/* ---- plane-fit objectives of the calibration refinement -------------------
 * No counterpart in the reference (the refinement is this library's). Per
 * view a set of points: xyz (n x 3 f64, row-major) and a selection byte per
 * point (a point counts where it is 1: a boolean mask, or the triangulation's
 * conditioning codes, 1 = conditioned); device pointers, on the context's
 * stream. The per-view results and parameters are host arrays; every sum is
 * FP64 in a fixed order (the same bits on every run), and each call returns
 * when its results are written.
 *   moments: count[v], mean[3v..], scatter[9v..] (the 3 x 3 matrix of the
 *     selected points about their mean, row-major);
 *   residuals: sumsq[v] = sum of ((x - mean) . normal)^2, sum[v] = sum of
 *     (x - mean) . normal over the selected points;
 *   grad: grad_xyz = a (x - mean) + b at selected points (a: 3 x 3 row-major,
 *     b: 3 per view), 0 elsewhere; the gradient of an objective of the
 *     scatter matrix S with dL/dS = G is a = G + G^T, b = 0 (the mean's
 *     terms sum to zero over the selected points). */
typedef struct psgc_plane_view {
    const double* xyz;
    const uint8_t* sel;
    int64_t n;
    double* grad_xyz;         /* grad: n x 3 output                       */
} psgc_plane_view;
PSGC_API int psgc_plane_moments_batch(psgc_ctx* ctx, const psgc_plane_view* views, int n,
                                      double* count, double* mean, double* scatter);
PSGC_API int psgc_plane_residuals_batch(psgc_ctx* ctx, const psgc_plane_view* views, int n,
                                        const double* mean, const double* normal, double* sumsq,
                                        double* sum);
PSGC_API int psgc_plane_grad_batch(psgc_ctx* ctx, const psgc_plane_view* views, int n,
                                   const double* mean, const double* a, const double* b);

/* ---- instrumentation ------------------------------------------------------
 * Per-kernel-family timing with hipEvents recorded on the launch stream.
 * psgc_profile_enable(1) clears and starts recording; psgc_profile_read
 * synchronises the device and returns, for family `name`, the number of
 * launches and their summed duration in milliseconds (0 launches if unknown). */
PSGC_API int psgc_profile_enable(int on);
PSGC_API int psgc_profile_read(const char* name, int64_t* launches, double* total_ms);

/* VALU issue ceilings of the context's device (no counterpart in the
 * reference; the measure the instruction-bound filter kernels are set
 * against): every wave of a grid that fills the device runs eight independent
 * fused multiply-add chains per lane, FP32 then FP64, no packed instructions.
 * Returns wave-instructions (64 lanes each) per second, in units of 1e9, of
 * one timed launch each after an untimed one. */
PSGC_API int psgc_probe_valu_rate(psgc_ctx* ctx, double* fp32_ginst_s, double* fp64_ginst_s);

#ifdef __cplusplus
}
#endif
#endif /* PSGC_H */
