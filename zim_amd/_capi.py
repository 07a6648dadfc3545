"""ctypes mirror of include/psgc.h (structs, enums) and the library loader."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsgc.so")

PSGC_OK = 0
PSGC_E_INVALID = -1
PSGC_E_NODEVICE = -2
PSGC_E_ALLOC = -3
PSGC_E_HIP = -4
PSGC_E_UNSUPPORTED = -5

MEM_DEVICE = 0
MEM_HOST = 1

SAMPLE_U8 = 0
SAMPLE_U16 = 1
SAMPLE_F64 = 2

UNWRAP_HETERODYNE3 = 0
UNWRAP_GRAYCODE = 1

SEM_CPU = 0
SEM_GPU = 1

OUTLIER_NONE = 0
OUTLIER_ROR = 1
OUTLIER_SOR = 2


class DecodeConfig(C.Structure):
    _fields_ = [
        ("freq", C.c_double * 3),
        ("shift_num", C.c_int32 * 3),
        ("modulation_threshold", C.c_double),
        ("preblur_sigma", C.c_double),
        ("avg_three_phase", C.c_int32),
        ("post_smooth", C.c_int32),
        ("enable_hdr", C.c_int32),
        ("hdr_saturation_gray", C.c_double),
        ("hdr_min_amplitude", C.c_double),
        ("unwrap_mode", C.c_int32),
        ("gray_bits", C.c_int32),
        ("gray_period_px", C.c_double),
        ("gray_min_contrast", C.c_double),
    ]

    @classmethod
    def default(cls, **kw) -> "DecodeConfig":
        """Defaults of JSLAB_PhaseShiftDecoder::Config
        (JSLAB_PhaseShiftDecoder.h:22-54); freq/shift_num have no default
        there and are set here to the reference driver's values
        (main.cpp:422-423)."""
        c = cls()
        c.freq[:] = (121.0, 110.0, 120.0)
        c.shift_num[:] = (8, 8, 8)
        c.modulation_threshold = 1.0
        c.preblur_sigma = 3.0
        c.avg_three_phase = 1
        c.post_smooth = 1
        c.enable_hdr = 0
        c.hdr_saturation_gray = 255.0
        c.hdr_min_amplitude = 2.0
        c.unwrap_mode = UNWRAP_HETERODYNE3
        c.gray_bits = 0
        c.gray_period_px = 0.0
        c.gray_min_contrast = 0.0
        for k, v in kw.items():
            if k in ("freq", "shift_num"):
                getattr(c, k)[:] = tuple(v)
            else:
                setattr(c, k, v)
        return c


class ReconConfig(C.Structure):
    _fields_ = [
        ("min_depth", C.c_double),
        ("max_depth", C.c_double),
        ("min_phase_diff", C.c_double),
        ("max_phase_diff", C.c_double),
        ("grad_threshold", C.c_double),
        ("use_rectified_coords", C.c_int32),
        ("use_disp_median_filter", C.c_int32),
        ("use_full_left_rectification", C.c_int32),
        ("radius_search", C.c_float),
        ("radius_min_neighbors", C.c_int32),
        ("radius_window_size", C.c_int32),
        ("use_mls", C.c_int32),
        ("mls_window_size", C.c_int32),
        ("mls_spatial_sigma", C.c_float),
        ("use_gpu_mls", C.c_int32),
        ("mls_polynomial_order", C.c_int32),
        ("mls_max_radius", C.c_float),
        ("mls_iterations", C.c_int32),
        ("semantics", C.c_int32),
        ("gauss_window", C.c_int32),
        ("gauss_sigma", C.c_float),
        ("outlier_mode", C.c_int32),
        ("sor_k", C.c_int32),
        ("sor_std_mul", C.c_float),
        ("sor_window", C.c_int32),
    ]

    @classmethod
    def default(cls, **kw) -> "ReconConfig":
        """Defaults of JSLAB_PhaseEpipolarReconstructor_V2::Config
        (JSLAB_PhaseEpipolarReconstructor_V2.h:18-55)."""
        c = cls(min_depth=0.0, max_depth=1000.0, min_phase_diff=0.03,
                max_phase_diff=0.4, grad_threshold=1.0, use_rectified_coords=1,
                use_disp_median_filter=1, use_full_left_rectification=1,
                radius_search=0.7, radius_min_neighbors=10, radius_window_size=10,
                use_mls=0, mls_window_size=30, mls_spatial_sigma=0.6, use_gpu_mls=1,
                mls_polynomial_order=1, mls_max_radius=0.0, mls_iterations=3,
                semantics=SEM_CPU, gauss_window=3, gauss_sigma=0.7,
                outlier_mode=OUTLIER_ROR, sor_k=30, sor_std_mul=1.0, sor_window=4)
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    @classmethod
    def default_gpu(cls, **kw) -> "ReconConfig":
        """Defaults of JSLAB_PhaseEpipolarReconstructor_V2_GPU::Config
        (JSLAB_PhaseEpipolarReconstructor_V2_GPU.h:27-61)."""
        c = cls.default(radius_search=0.05, radius_min_neighbors=5,
                        radius_window_size=5, use_mls=1, mls_window_size=25,
                        mls_spatial_sigma=0.5, mls_iterations=2, semantics=SEM_GPU)
        for k, v in kw.items():
            setattr(c, k, v)
        return c


KERNELS_TILED = 0
KERNELS_PLAIN = 1

BATCH_GRAPH = 1
BATCH_AUTOTUNE = 2


class PairDesc(C.Structure):
    _fields_ = [
        ("left", C.c_void_p * 64),
        ("right", C.c_void_p * 64),
        ("n_frames", C.c_int32),
        ("sample_type", C.c_int32),
        ("row_stride_bytes", C.c_int64),
        ("rows", C.c_int32),
        ("cols", C.c_int32),
        ("decode", DecodeConfig),
        ("recon", ReconConfig),
        ("pha_f", C.c_int32),
        ("Q", C.c_double * 16),
        ("R1", C.c_double * 9),
        ("xyz", C.c_void_p),
        ("P", C.c_void_p),
        ("disp", C.c_void_p),
        ("coef", C.c_void_p * 3),
    ]


class PlaneView(C.Structure):
    """psgc_plane_view (include/psgc.h)."""
    _fields_ = [
        ("xyz", C.c_void_p),
        ("sel", C.c_void_p),
        ("n", C.c_int64),
        ("grad_xyz", C.c_void_p),
    ]


# hipStreamLegacy: the legacy default stream, addressed explicitly
HIP_STREAM_LEGACY = 1


def stream_handle(stream) -> int:
    """The hipStream_t to give the library for a torch stream. torch's
    default stream is the legacy null stream (handle 0), which the library
    would read as "the context's own stream": it gets hipStreamLegacy
    instead, so that library work is ordered with torch's and with every
    other context's on that stream."""
    return stream.cuda_stream or HIP_STREAM_LEGACY


def ptr(x):
    """Address of a numpy array, a torch tensor, or None."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return C.c_void_p(x.data_ptr())
    return C.c_void_p(x.ctypes.data)


_lib = None


def load(path: str | None = None):
    """Loads libpsgc.so (import torch first so that the process has a single
    HIP runtime: the library resolves libamdhip64.so.7 to torch's copy)."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (one HIP runtime per process)
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built: run __graft_entry__.build() (the HIP path "
            "has no CPU fallback)")
    lib = C.CDLL(path)
    vp, i32, i64, dbl, flt = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_float
    sig = {
        "psgc_last_error": ([], C.c_char_p),
        "psgc_abi_version": ([], i32),
        "psgc_device_count": ([], i32),
        "psgc_create": ([C.POINTER(vp), i32], i32),
        "psgc_destroy": ([vp], i32),
        "psgc_set_stream": ([vp, vp], i32),
        "psgc_synchronize": ([vp], i32),
        "psgc_warmup": ([], i32),
        "psgc_release": ([vp], i32),
        "psgc_decode_config_default": ([vp], None),
        "psgc_recon_config_default": ([vp], None),
        "psgc_recon_config_default_gpu": ([vp], None),
        "psgc_decode": ([vp, vp, vp, i32, i32, i32, i32, i64, vp, vp, vp, vp, i32], i32),
        "psgc_reconstruct": ([vp, vp, vp, vp, vp, vp, i32, i32, i32, vp, vp, vp, vp, vp,
                              vp, i32], i32),
        "psgc_stereo_match": ([vp, vp, vp, vp, vp, i32, i32, i32, dbl, vp, vp, i32], i32),
        "psgc_median5": ([vp, vp, vp, i32, i32, i32], i32),
        "psgc_reproject": ([vp, vp, i32, i32, vp, vp, dbl, dbl, i32, vp, vp, i32], i32),
        "psgc_reproject_poly": ([vp, vp, i32, i32, vp, vp, vp, vp, dbl, dbl, i32, vp, vp, i32],
                                i32),
        "psgc_reconstruct_poly": ([vp, vp, vp, vp, vp, vp, i32, i32, i32, vp, vp, vp, vp, vp,
                                   vp, vp, vp, i32], i32),
        "psgc_load_coef_bin": ([C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), vp], i32),
        "psgc_roi_right_window": ([vp, vp, vp, i32, i32, i32, vp, vp, vp, i32], i32),
        "psgc_reconstruct_poly_roi": ([vp, vp, vp, vp, vp, vp, i32, i32, i32, vp, vp, vp, vp,
                                       vp, vp, vp, vp, vp, vp, i32], i32),
        "psgc_filter_gaussian": ([vp, vp, vp, vp, vp, i32, i32, i32, flt, i32, i32], i32),
        "psgc_filter_ror": ([vp, vp, vp, vp, vp, i32, i32, flt, i32, i32, i32], i32),
        "psgc_filter_sor": ([vp, vp, vp, vp, vp, i32, i32, i32, flt, i32, vp, vp, i32], i32),
        "psgc_cloud_sor": ([vp, vp, vp, C.c_int64, i32, flt, flt, i32, vp, vp, vp, i32], i32),
        "psgc_cloud_tune_voxel": ([vp, vp, vp, C.c_int64, i32, i32, vp, i32, vp, vp], i32),
        "psgc_cloud_mls": ([vp, vp, vp, C.c_int64, flt, i32, flt, i32, i32, flt, vp, i32], i32),
        "psgc_cloud_gaussian": ([vp, vp, vp, C.c_int64, flt, flt, flt, vp, i32], i32),
        "psgc_filter_mls": ([vp, vp, vp, vp, i32, i32, i32, flt, i32, flt, i32, i32, i32],
                            i32),
        "psgc_rectify_ideal": ([dbl, dbl, dbl, dbl, dbl, dbl, vp, vp], i32),
        "psgc_full_left_rectification": ([vp, i32, i32, vp, vp, vp, vp, vp, vp], i32),
        "psgc_stereo_rectify": ([vp, vp, vp, vp, i32, i32, vp, vp, vp, vp, vp], i32),
        "psgc_profile_enable": ([i32], i32),
        "psgc_set_kernels": ([vp, i32], i32),
        "psgc_rectify_maps": ([vp, vp, vp, vp, i32, i32, vp, vp, i32], i32),
        "psgc_remap": ([vp, i32, vp, i32, i32, vp, vp, i32, i32, vp, i32], i32),
        "psgc_compact": ([vp, vp, vp, vp, i32, i32, vp, vp, C.POINTER(C.c_int64), i32], i32),
        "psgc_compact_noalloc": ([vp, vp, vp, vp, i32, i32, vp, vp, vp], i32),
        "psgc_write_ply": ([C.c_char_p, vp, vp, C.c_int64], i32),
        "psgc_write_raw_f64": ([C.c_char_p, C.c_char_p, vp, i32, i32], i32),
        "psgc_adjoint_forward": ([vp, vp, vp, vp, vp, vp, i32, i32, vp, vp, i32], i32),
        "psgc_adjoint_backward": ([vp, vp, vp, vp, vp, vp, i32, i32, vp, vp, vp, vp, i32], i32),
        "psgc_batch_create": ([vp, vp, i32, i32, C.POINTER(vp)], i32),
        "psgc_batch_run": ([vp], i32),
        "psgc_batch_run_eager": ([vp], i32),
        "psgc_batch_destroy": ([vp], i32),
        "psgc_batch_grid": ([vp, C.c_char_p, C.POINTER(C.c_int)], i32),
        "psgc_batch_set_grid": ([vp, C.c_char_p, i32], i32),
        "psgc_profile_read": ([C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)], i32),
        "psgc_probe_valu_rate": ([C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)], i32),
        "psgc_plane_moments_batch": ([vp, vp, i32, vp, vp, vp], i32),
        "psgc_plane_residuals_batch": ([vp, vp, i32, vp, vp, vp, vp], i32),
        "psgc_plane_grad_batch": ([vp, vp, i32, vp, vp, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


class PsgcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"psgc error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != PSGC_OK:
        msg = _lib.psgc_last_error().decode() if _lib is not None else ""
        raise PsgcError(rc, msg)
    return rc
