"""TEST INFRASTRUCTURE: ctypes binding of the CPU oracle (libpsgc_oracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, always as the checker (or the timed CPU baseline), never as part of
the measured or shipped path. Build it with `make -C oracle`.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from zim_amd._capi import DecodeConfig, ReconConfig  # struct layouts only

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsgc_oracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: make -C oracle")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, dbl, flt = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_float
        L.orc_decode.argtypes = [vp, vp, i32, i32, i32, i32, i64, vp, vp, vp, vp]
        L.orc_stereo_match.argtypes = [vp, vp, vp, vp, i32, i32, i32, dbl, vp, vp]
        L.orc_median5.argtypes = [vp, vp, i32, i32]
        L.orc_reproject.argtypes = [vp, i32, i32, vp, vp, dbl, dbl, i32, vp, vp]
        L.orc_gaussian.argtypes = [vp, vp, vp, vp, i32, i32, i32, flt, i32]
        L.orc_ror.argtypes = [vp, vp, vp, vp, i32, i32, flt, i32, i32]
        L.orc_sor.argtypes = [vp, vp, vp, vp, i32, i32, i32, flt, i32, vp, vp]
        L.orc_mls.argtypes = [vp, vp, vp, i32, i32, i32, flt, i32, flt, i32, i32]
        L.orc_eigen3.argtypes = [vp, vp, vp, i32]
        L.orc_chol.argtypes = [vp, vp, vp, i32]
        L.orc_reconstruct.argtypes = [vp, vp, vp, vp, vp, i32, i32, i32, vp, vp, vp, vp,
                                      vp, vp]
        L.orc_atan2_many.argtypes = [vp, vp, vp, C.c_long]
        L.orc_remap.argtypes = [i32, vp, i32, i32, vp, vp, i32, i32, vp]
        L.orc_voxel_cells.argtypes = [vp, i64, flt, flt, vp]
        L.orc_knn_point_limit.argtypes = [vp, i64, vp, flt, vp]
        L.orc_reproject_poly.argtypes = [vp, i32, i32, vp, vp, vp, vp, dbl, dbl, i32, vp, vp]
        L.orc_roi_right_window.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp]
        L.orc_reconstruct_poly_roi.argtypes = [vp, vp, vp, vp, vp, i32, i32, i32, vp, vp, vp,
                                               vp, vp, vp, vp, vp, vp, vp]
        L.orc_reconstruct_poly.argtypes = [vp, vp, vp, vp, vp, i32, i32, i32, vp, vp, vp, vp,
                                           vp, vp, vp, vp]
        L.orc_num_threads.restype = i32
        L.orc_check_f32.argtypes = [i32, flt, flt, i32, vp, vp, vp, vp]
        L.orc_check_exp_w.argtypes = [flt, flt, i32, vp, vp]
        L.orc_check_sort_net.argtypes = [i32, vp]
        i64 = C.c_int64
        L.orc_cloud_sor.argtypes = [vp, vp, i64, i32, flt, flt, i32, vp, vp, vp]
        L.orc_cloud_mls.argtypes = [vp, vp, i64, flt, i32, flt, i32, i32, flt, vp]
        L.orc_cloud_gaussian.argtypes = [vp, vp, i64, flt, flt, flt, vp]
        L.orc_cloud_brute.argtypes = [vp, vp, i64, i32, flt, vp, vp]
        _lib = L
    return _lib


def check_f32(which: str, lo: float, hi: float, stride: int = 1):
    """Compares psgc::{exp,acos,cos}_f32 with the C library over every
    `stride`-th float in [lo, hi]; returns (max ulp vs the float library
    routine, #differences from it, #differences from the double routine
    rounded to float, #floats tested)."""
    out = [C.c_int64() for _ in range(4)]
    rc = lib().orc_check_f32({"exp": 0, "acos": 1, "cos": 2}[which], lo, hi, stride,
                             *[C.byref(o) for o in out])
    assert rc == 0
    return tuple(o.value for o in out)


def check_exp_w(lo: float, hi: float, stride: int = 1):
    """Floats in [lo, hi] (every `stride`-th) where psgc::exp_f32_w, the
    organized MLS kernel's weight exponential, differs from psgc::exp_f32 in
    any bit; returns (#differences, #floats tested)."""
    nd, nt = C.c_int64(), C.c_int64()
    assert lib().orc_check_exp_w(lo, hi, stride, C.byref(nd), C.byref(nt)) == 0
    return nd.value, nt.value


def check_sort_net(k: int) -> int:
    """Inputs of zeros and ones (all 2^k) that the device's k-input sorting
    network leaves unsorted."""
    out = C.c_int64()
    assert lib().orc_check_sort_net(k, C.byref(out)) == 0
    return out.value


def _pts(xyz, valid):
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    v = None if valid is None else np.ascontiguousarray(valid, np.uint8).reshape(-1)
    return xyz, v


def cloud_sor(xyz, valid, k, std_mul, voxel, rmax=4):
    """SOR on an unorganized cloud (N x 3 f32): (keep u8, mean distance f64,
    stats {count, mean, sd, threshold})."""
    xyz, v = _pts(xyz, valid)
    n = len(xyz)
    keep = np.zeros(n, np.uint8)
    md = np.zeros(n, np.float64)
    st = np.zeros(4, np.float64)
    rc = lib().orc_cloud_sor(_p(xyz), _p(v), n, k, std_mul, voxel, rmax, _p(keep), _p(md), _p(st))
    assert rc == 0, rc
    return keep, md, st


def cloud_mls(xyz, valid, sigma, order, max_radius, iterations, semantics, voxel):
    xyz, v = _pts(xyz, valid)
    out = np.empty_like(xyz)
    rc = lib().orc_cloud_mls(_p(xyz), _p(v), len(xyz), sigma, order, max_radius, iterations,
                             semantics, voxel, _p(out))
    assert rc == 0, rc
    return out


def cloud_gaussian(xyz, valid, sigma, radius, voxel):
    xyz, v = _pts(xyz, valid)
    out = np.empty_like(xyz)
    rc = lib().orc_cloud_gaussian(_p(xyz), _p(v), len(xyz), sigma, radius, voxel, _p(out))
    assert rc == 0, rc
    return out


def cloud_brute(xyz, valid, k, radius):
    """O(N^2) pins: k smallest squared distances to other points (+inf pads)
    and the number of points within radius (self included)."""
    xyz, v = _pts(xyz, valid)
    n = len(xyz)
    d2 = np.empty((n, k), np.float32)
    cnt = np.empty(n, np.int32)
    assert lib().orc_cloud_brute(_p(xyz), _p(v), n, k, radius, _p(d2), _p(cnt)) == 0
    return d2, cnt


def num_threads() -> int:
    """OpenMP threads the oracle's parallel loops use."""
    return int(lib().orc_num_threads())


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


SAMPLE = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float64): 2}


def decode(cfg: DecodeConfig, frames):
    """frames: list of 2-D arrays (u8/u16/f64). Returns phase, k, mod, valid."""
    frames = [np.ascontiguousarray(f) for f in frames]
    rows, cols = frames[0].shape
    arr = (C.c_void_p * len(frames))(*[f.ctypes.data for f in frames])
    ph = np.zeros((rows, cols), np.float64)
    k = np.zeros((rows, cols), np.int32)
    md = np.zeros((rows, cols), np.float64)
    vd = np.zeros((rows, cols), np.uint8)
    rc = lib().orc_decode(C.byref(cfg), arr, len(frames), SAMPLE[frames[0].dtype], rows, cols,
                          0, _p(ph), _p(k), _p(md), _p(vd))
    if rc != 0:
        raise ValueError(f"orc_decode rc={rc}")
    return ph, k, md, vd


def stereo_match(lp, rp, lk, rk, pha_f, grad_threshold=1.0):
    lp, rp = _c(lp, np.float64), _c(rp, np.float64)
    lk, rk = _c(lk, np.int32), _c(rk, np.int32)
    rows, cols = lp.shape
    disp = np.zeros((rows, cols), np.float64)
    mc = np.zeros((rows, cols), np.int32)
    lib().orc_stereo_match(_p(lp), _p(rp), _p(lk), _p(rk), rows, cols, pha_f, grad_threshold,
                           _p(disp), _p(mc))
    return disp, mc


def median5(src):
    src = _c(src, np.float32)
    dst = np.empty_like(src)
    lib().orc_median5(_p(src), _p(dst), *src.shape)
    return dst


def reproject(disp, Q, R1, min_depth=0.0, max_depth=1000.0, use_rect=1):
    disp = _c(disp, np.float64)
    rows, cols = disp.shape
    xyz = np.zeros((rows, cols, 3), np.float32)
    P = np.zeros((rows, cols), np.uint8)
    Q, R1 = _c(Q, np.float64), _c(R1, np.float64)
    lib().orc_reproject(_p(disp), rows, cols, _p(Q), _p(R1), min_depth, max_depth, use_rect,
                        _p(xyz), _p(P))
    return xyz, P


def gaussian(xyz, P, win=3, sigma=0.7, semantics=0):
    xyz, P = _c(xyz, np.float32), _c(P, np.uint8)
    rows, cols = P.shape
    d, dP = np.zeros_like(xyz), np.zeros_like(P)
    lib().orc_gaussian(_p(xyz), _p(P), _p(d), _p(dP), rows, cols, win, sigma, semantics)
    return d, dP


def ror(xyz, P, radius, min_neighbors, window):
    xyz, P = _c(xyz, np.float32), _c(P, np.uint8)
    rows, cols = P.shape
    d, dP = np.zeros_like(xyz), np.zeros_like(P)
    lib().orc_ror(_p(xyz), _p(P), _p(d), _p(dP), rows, cols, radius, min_neighbors, window)
    return d, dP


def sor(xyz, P, k, std_mul, half_window):
    xyz, P = _c(xyz, np.float32), _c(P, np.uint8)
    rows, cols = P.shape
    d, dP = np.zeros_like(xyz), np.zeros_like(P)
    md = np.zeros((rows, cols), np.float64)
    st = np.zeros(4, np.float64)
    lib().orc_sor(_p(xyz), _p(P), _p(d), _p(dP), rows, cols, k, std_mul, half_window, _p(md),
                  _p(st))
    return d, dP, md, st


def mls(xyz, P, window, sigma, order, max_radius=0.0, iterations=1, semantics=0):
    xyz, P = _c(xyz, np.float32), _c(P, np.uint8)
    rows, cols = P.shape
    d = np.zeros_like(xyz)
    lib().orc_mls(_p(xyz), _p(P), _p(d), rows, cols, window, sigma, order, max_radius,
                  iterations, semantics)
    return d


def eigen3(C):
    """The MLS filter's symmetric 3x3 eigen solver on a stack of float32
    matrices (..., 3, 3): eigenvalues (..., 3) in the solver's ascending
    order and eigenvectors (..., 3, 3), one per row."""
    C = np.ascontiguousarray(C, np.float32)
    shp = C.shape[:-2]
    Cf = C.reshape(-1, 3, 3)
    ev = np.empty((len(Cf), 3), np.float32)
    V = np.empty((len(Cf), 3, 3), np.float32)
    assert lib().orc_eigen3(_p(Cf), _p(ev), _p(V), len(Cf)) == 0
    return ev.reshape(*shp, 3), V.reshape(*shp, 3, 3)


def chol(A, b):
    """The MLS filter's float Cholesky solve of A x = b (n <= 6): (x, ok)."""
    A = np.ascontiguousarray(A, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    n = len(b)
    x = np.zeros(n, np.float32)
    rc = lib().orc_chol(_p(A), _p(b), _p(x), n)
    assert rc in (0, 1), rc
    return x, bool(rc)


def reconstruct(cfg: ReconConfig, lp, rp, lk, rk, pha_f, Q, R1):
    lp, rp = _c(lp, np.float64), _c(rp, np.float64)
    lk, rk = _c(lk, np.int32), _c(rk, np.int32)
    rows, cols = lp.shape
    xyz = np.zeros((rows, cols, 3), np.float32)
    P = np.zeros((rows, cols), np.uint8)
    disp = np.zeros((rows, cols), np.float32)
    mc = np.zeros((rows, cols), np.int32)
    Q, R1 = _c(Q, np.float64), _c(R1, np.float64)
    lib().orc_reconstruct(C.byref(cfg), _p(lp), _p(rp), _p(lk), _p(rk), rows, cols, pha_f,
                          _p(Q), _p(R1), _p(xyz), _p(P), _p(disp), _p(mc))
    return xyz, P, disp, mc


def reproject_poly(disp, coef, R1, min_depth=0.0, max_depth=1000.0, use_rect=1):
    """coef: (cx, cy, cz), each rows x cols x 4 float32 (the _rx/_ry/_rz maps)."""
    disp = _c(disp, np.float64)
    rows, cols = disp.shape
    cx, cy, cz = [_c(c, np.float32) for c in coef]
    for c in (cx, cy, cz):
        assert c.shape == (rows, cols, 4)
    xyz = np.zeros((rows, cols, 3), np.float32)
    P = np.zeros((rows, cols), np.uint8)
    R1 = _c(R1, np.float64)
    lib().orc_reproject_poly(_p(disp), rows, cols, _p(cx), _p(cy), _p(cz), _p(R1), min_depth,
                             max_depth, use_rect, _p(xyz), _p(P))
    return xyz, P


def reconstruct_poly(cfg: ReconConfig, lp, rp, lk, rk, pha_f, coef, R1):
    lp, rp = _c(lp, np.float64), _c(rp, np.float64)
    lk, rk = _c(lk, np.int32), _c(rk, np.int32)
    rows, cols = lp.shape
    cx, cy, cz = [_c(c, np.float32) for c in coef]
    for c in (cx, cy, cz):
        assert c.shape == (rows, cols, 4)
    xyz = np.zeros((rows, cols, 3), np.float32)
    P = np.zeros((rows, cols), np.uint8)
    disp = np.zeros((rows, cols), np.float32)
    mc = np.zeros((rows, cols), np.int32)
    R1 = _c(R1, np.float64)
    lib().orc_reconstruct_poly(C.byref(cfg), _p(lp), _p(rp), _p(lk), _p(rk), rows, cols, pha_f,
                               _p(cx), _p(cy), _p(cz), _p(R1), _p(xyz), _p(P), _p(disp), _p(mc))
    return xyz, P, disp, mc


def roi_right_window(lk, rk, pha_f, left_rect, mask):
    lk, rk = _c(lk, np.int32), _c(rk, np.int32)
    rows, cols = lk.shape
    lr = np.asarray(left_rect, np.int32)
    mask = _c(mask, np.uint8)
    assert mask.shape == (lr[3], lr[2])
    rr = np.zeros(4, np.int32)
    lib().orc_roi_right_window(_p(lk), _p(rk), rows, cols, pha_f, _p(lr), _p(mask), _p(rr))
    return rr


def reconstruct_poly_roi(cfg: ReconConfig, lp, rp, lk, rk, pha_f, left_rect, right_rect, mask,
                         coef, R1):
    """startCalculate_v2_roi; outputs have the left rectangle's size."""
    lp, rp = _c(lp, np.float64), _c(rp, np.float64)
    lk, rk = _c(lk, np.int32), _c(rk, np.int32)
    rows, cols = lp.shape
    lr, rr = np.asarray(left_rect, np.int32), np.asarray(right_rect, np.int32)
    mask = _c(mask, np.uint8)
    assert mask.shape == (lr[3], lr[2])
    cx, cy, cz = [_c(c, np.float32) for c in coef]
    xyz = np.zeros((lr[3], lr[2], 3), np.float32)
    P = np.zeros((lr[3], lr[2]), np.uint8)
    disp = np.zeros((lr[3], lr[2]), np.float32)
    R1 = _c(R1, np.float64)
    lib().orc_reconstruct_poly_roi(C.byref(cfg), _p(lp), _p(rp), _p(lk), _p(rk), rows, cols,
                                   pha_f, _p(lr), _p(rr), _p(mask), _p(cx), _p(cy), _p(cz),
                                   _p(R1), _p(xyz), _p(P), _p(disp))
    return xyz, P, disp


def voxel_cells(v, o, voxel):
    """The voxel frame's cell of each coordinate in v (origin o, edge voxel)."""
    v = _c(v, np.float32)
    out = np.zeros(v.shape, np.int64)
    assert lib().orc_voxel_cells(_p(v), v.size, float(o), float(voxel), _p(out)) == 0
    return out


def knn_point_limit(xyz, o, voxel):
    """The device's per-point bound on the squared k-th distance that its
    3x3x3 k-nearest pass settles (psgc_voxel.h, cloud_knn_point_limit)."""
    xyz = _c(xyz, np.float32).reshape(-1, 3)
    o = _c(o, np.float32)
    out = np.zeros(len(xyz), np.float32)
    assert lib().orc_knn_point_limit(_p(xyz), len(xyz), _p(o), float(voxel), _p(out)) == 0
    return out


def remap(src, map_x, map_y, nearest=False):
    """cv::remap (OpenCV 4.7.0) of a f64 map with INTER_LINEAR, or of an i32
    map with INTER_NEAREST, BORDER_CONSTANT 0."""
    src = _c(src, np.int32 if nearest else np.float64)
    mx, my = _c(map_x, np.float32), _c(map_y, np.float32)
    assert mx.shape == my.shape and src.ndim == 2
    dst = np.zeros(mx.shape, src.dtype)
    r = lib().orc_remap(1 if nearest else 0, _p(src), src.shape[0], src.shape[1], _p(mx), _p(my),
                        mx.shape[0], mx.shape[1], _p(dst))
    assert r == 0
    return dst


def atan2(y, x):
    y, x = _c(y, np.float64), _c(x, np.float64)
    out = np.empty_like(y)
    lib().orc_atan2_many(_p(y), _p(x), _p(out), y.size)
    return out
