"""zim_amd — MI355X-native reconstruction path of PhaseShiftGrayCode3D.

Python host side over the C ABI of libpsgc.so (include/psgc.h). It mirrors
the reference's operator interface for this path:

  PhaseShiftDecoder.create(config, Version.V2)      JSLAB_PhaseShiftDecoder.h:10-92
    .setImages(frames) / .startCalculate() / .getOrderMap() / .getModulationMap()
  PhaseEpipolarReconstructor_V2(config)             JSLAB_PhaseEpipolarReconstructor_V2.h:11-171
    .setCalibrationParams(...) / .setData(...) / .startCalculate() / .getDisparityMap()
  applyMLSFilterToPointCloud(...)                   JSLAB_MLS_Filter.h:53-57

Device tensors (torch, on the GPU) are passed to the library as HBM
pointers and the work is queued on torch's current stream; numpy arrays take
the host-staging path of the C ABI. There is no CPU fallback: without the
built library or a GPU every call raises.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import enum
import math

import numpy as np
import torch

from . import _capi
from ._capi import (DecodeConfig, ReconConfig, PsgcError, MEM_DEVICE, MEM_HOST,  # noqa: F401
                    SEM_CPU, SEM_GPU, OUTLIER_NONE, OUTLIER_ROR, OUTLIER_SOR,
                    UNWRAP_HETERODYNE3, UNWRAP_GRAYCODE)

__all__ = ["Batch", "Context", "DecodeConfig", "ReconConfig", "PhaseShiftDecoder", "Version",
           "PhaseEpipolarReconstructor_V2", "applyMLSFilterToPointCloud", "decode",
           "reconstruct", "PsgcError", "full_left_rectification", "rectify_maps", "remap",
           "stereo_rectify", "StereoCalibParams_V1"]

_SAMPLE = {torch.uint8: _capi.SAMPLE_U8, torch.uint16: _capi.SAMPLE_U16,
           torch.float64: _capi.SAMPLE_F64, np.dtype(np.uint8): _capi.SAMPLE_U8,
           np.dtype(np.uint16): _capi.SAMPLE_U16, np.dtype(np.float64): _capi.SAMPLE_F64}


class Context:
    """Owns a psgc_ctx on one device; its work runs on torch's current stream."""

    def __init__(self, device: int = 0):
        self.lib = _capi.load()
        self.device = device
        h = C.c_void_p()
        _capi.check(self.lib.psgc_create(C.byref(h), device))
        self.h = h
        self._bound_stream = None

    def set_kernels(self, which: str):
        """"tiled" (LDS-tiled kernels, the default) or "plain" (one thread
        per pixel, the reference GPU path's structure)."""
        code = {"tiled": _capi.KERNELS_TILED, "plain": _capi.KERNELS_PLAIN}[which]
        _capi.check(self.lib.psgc_set_kernels(self.h, code))
        return self

    def bind_stream(self):
        h = _capi.stream_handle(torch.cuda.current_stream(self.device))
        if self._bound_stream != h:
            _capi.check(self.lib.psgc_set_stream(self.h, C.c_void_p(h)))
            self._bound_stream = h
        return self

    def release(self):
        _capi.check(self.lib.psgc_release(self.h))

    def close(self):
        if self.h:
            self.lib.psgc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = {}


def default_context(device: int = 0) -> Context:
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device].bind_stream()


def mls_half_window(radius, f, z_min):
    """First estimate of the half window that holds an MLS radius (mm) for a
    pair whose nearest surface lies at z_min (mm), for focal length f (px):
    r f / z_min pixels span the radius on a fronto-parallel surface, plus a
    half-pixel margin. Points off the optical axis, tilted surfaces and
    smoothing can need more: a window holds the radius for a cloud when its
    outer ring has no neighbour within the radius, which the caller checks on
    the cloud itself. The organized window of the reference's filter
    (JSLAB_MLS_Filter.cpp:191-194) is otherwise a cut through the ball."""
    return int(math.ceil(radius * f / z_min + 0.5))


def _is_dev(x):
    return isinstance(x, torch.Tensor) and x.is_cuda


_NP = {torch.float64: np.float64, torch.float32: np.float32, torch.int32: np.int32,
       torch.uint8: np.uint8}


def _map(name, a, dev, device, dtype, shape):
    """One input map of a call, checked before its address reaches a kernel:
    device tensors must sit on the call's device with exactly `dtype`, `shape`
    and a dense layout (a wrong one would be read out of bounds or as the
    wrong type); host arrays must have `shape` and are converted to `dtype`."""
    if dev:
        if not _is_dev(a):
            raise ValueError(f"{name}: expected a device tensor like the call's other inputs")
        if a.device != device:
            raise ValueError(f"{name}: on {a.device}, the call's other inputs on {device}")
        if a.dtype != dtype:
            raise ValueError(f"{name}: dtype {a.dtype}, expected {dtype}")
        if tuple(a.shape) != tuple(shape):
            raise ValueError(f"{name}: shape {tuple(a.shape)}, expected {tuple(shape)}")
        if not a.is_contiguous():
            raise ValueError(f"{name}: must be contiguous")
        return a
    if _is_dev(a):
        raise ValueError(f"{name}: expected a host array like the call's other inputs")
    arr = np.ascontiguousarray(a.numpy() if isinstance(a, torch.Tensor) else a, _NP[dtype])
    if arr.shape != tuple(shape):
        raise ValueError(f"{name}: shape {arr.shape}, expected {tuple(shape)}")
    return arr


def _frames(frames, name="frames"):
    """A camera's frame stack: one sample type, one shape; on the device one
    device and one row stride with unit column stride (the ABI takes one
    stride for all frames). Returns (frames, is_device, sample type, row
    stride bytes or 0, rows, cols)."""
    if len(frames) == 0:
        raise ValueError(f"{name}: no frames")
    f0 = frames[0]
    dev = _is_dev(f0)
    if f0.ndim != 2:
        raise ValueError(f"{name}: frames must be 2-D")
    if f0.dtype not in _SAMPLE:
        raise ValueError(f"{name}: sample type {f0.dtype} is not uint8, uint16 or float64")
    if dev:
        for i, f in enumerate(frames):
            if not _is_dev(f) or f.device != f0.device:
                raise ValueError(f"{name}[{i}]: not on {f0.device}")
            if f.dtype != f0.dtype or tuple(f.shape) != tuple(f0.shape):
                raise ValueError(f"{name}[{i}]: {f.dtype} {tuple(f.shape)}, frame 0 is "
                                 f"{f0.dtype} {tuple(f0.shape)}")
            if f.stride(-1) != 1 or f.stride(0) != f0.stride(0) or f.stride(0) < f0.shape[1]:
                raise ValueError(f"{name}[{i}]: rows must be dense with the stride of frame 0")
        stride = f0.stride(0) * f0.element_size()
    else:
        frames = [np.ascontiguousarray(f) for f in frames]
        for i, f in enumerate(frames):
            if f.dtype != frames[0].dtype or f.shape != frames[0].shape:
                raise ValueError(f"{name}[{i}]: {f.dtype} {f.shape}, frame 0 is "
                                 f"{frames[0].dtype} {frames[0].shape}")
        stride = 0
    rows, cols = frames[0].shape
    return frames, dev, _SAMPLE[frames[0].dtype], stride, rows, cols


def _out(like_dev, shape, dtype, device):
    if like_dev:
        return torch.empty(shape, dtype=dtype, device=device)
    return np.empty(shape, dtype=torch.empty(0, dtype=dtype).numpy().dtype)


def decode(cfg: DecodeConfig, frames, ctx: Context | None = None, want_mod=True,
           want_valid=True):
    """Decodes one camera's frame stack. frames: list of 2-D arrays/tensors of
    one sample type (u8, u16 or f64). Returns (phase f64, k i32, mod f64,
    valid u8)."""
    ctx = ctx or default_context()
    frames, dev, st, stride, rows, cols = _frames(frames)
    ptrs = [f.data_ptr() if dev else f.ctypes.data for f in frames]
    device = frames[0].device if dev else None
    arr = (C.c_void_p * len(ptrs))(*ptrs)
    ph = _out(dev, (rows, cols), torch.float64, device)
    k = _out(dev, (rows, cols), torch.int32, device)
    md = _out(dev, (rows, cols), torch.float64, device) if want_mod else None
    vd = _out(dev, (rows, cols), torch.uint8, device) if want_valid else None
    _capi.check(ctx.lib.psgc_decode(ctx.h, C.byref(cfg), arr, len(ptrs), st, rows, cols, stride,
                                    _capi.ptr(ph), _capi.ptr(k), _capi.ptr(md), _capi.ptr(vd),
                                    MEM_DEVICE if dev else MEM_HOST))
    return ph, k, md, vd


def _recon_maps(lp, rp, lk, rk):
    """The four rectified maps of a reconstruction: f64 phases and i32 orders
    of one shape, all on the host or all on one device."""
    dev = _is_dev(lp)
    if getattr(lp, "ndim", 0) != 2:
        raise ValueError("l_phase must be 2-D")
    rows, cols = lp.shape
    device = lp.device if dev else None
    maps = [_map(n, a, dev, device, t, (rows, cols))
            for n, a, t in (("l_phase", lp, torch.float64), ("r_phase", rp, torch.float64),
                            ("l_k", lk, torch.int32), ("r_k", rk, torch.int32))]
    return dev, device, rows, cols, maps


def reconstruct(cfg: ReconConfig, lp, rp, lk, rk, pha_f, Q, R1, ctx: Context | None = None):
    """Match -> median -> reproject -> Gaussian -> outlier -> MLS on rectified
    maps. Returns (xyz f32 HxWx3, P u8, disp f32, match_col i32)."""
    ctx = ctx or default_context()
    dev, device, rows, cols, (lp, rp, lk, rk) = _recon_maps(lp, rp, lk, rk)
    xyz = _out(dev, (rows, cols, 3), torch.float32, device)
    P = _out(dev, (rows, cols), torch.uint8, device)
    disp = _out(dev, (rows, cols), torch.float32, device)
    mc = _out(dev, (rows, cols), torch.int32, device)
    Qa = np.ascontiguousarray(Q, np.float64).reshape(16)
    Ra = np.ascontiguousarray(R1, np.float64).reshape(9)
    _capi.check(ctx.lib.psgc_reconstruct(ctx.h, C.byref(cfg), _capi.ptr(lp), _capi.ptr(rp),
                                         _capi.ptr(lk), _capi.ptr(rk), rows, cols, int(pha_f),
                                         _capi.ptr(Qa), _capi.ptr(Ra), _capi.ptr(xyz),
                                         _capi.ptr(P), _capi.ptr(disp), _capi.ptr(mc),
                                         MEM_DEVICE if dev else MEM_HOST))
    return xyz, P, disp, mc


def _coef_maps(coef, dev, rows, cols, device=None):
    if len(coef) != 3:
        raise ValueError("coef: expected the three maps (cx, cy, cz)")
    if dev and device is None:
        device = coef[0].device if _is_dev(coef[0]) else None
    return [_map(f"coef[{a}]", c, dev, device, torch.float32, (rows, cols, 4))
            for a, c in enumerate(coef)]


def reconstruct_poly(cfg: ReconConfig, lp, rp, lk, rk, pha_f, coef, R1,
                     ctx: Context | None = None):
    """startCalculate_v2: reconstruct with the per-pixel polynomial
    triangulation. coef = (cx, cy, cz), each rows x cols x 4 float32 (the
    _rx/_ry/_rz calibration maps). Returns (xyz, P, disp, match_col)."""
    ctx = ctx or default_context()
    dev, device, rows, cols, (lp, rp, lk, rk) = _recon_maps(lp, rp, lk, rk)
    xyz = _out(dev, (rows, cols, 3), torch.float32, device)
    P = _out(dev, (rows, cols), torch.uint8, device)
    disp = _out(dev, (rows, cols), torch.float32, device)
    mc = _out(dev, (rows, cols), torch.int32, device)
    cx, cy, cz = _coef_maps(coef, dev, rows, cols, device)
    Ra = np.ascontiguousarray(R1, np.float64).reshape(9)
    _capi.check(ctx.lib.psgc_reconstruct_poly(ctx.h, C.byref(cfg), _capi.ptr(lp), _capi.ptr(rp),
                                              _capi.ptr(lk), _capi.ptr(rk), rows, cols,
                                              int(pha_f), _capi.ptr(cx), _capi.ptr(cy),
                                              _capi.ptr(cz), _capi.ptr(Ra), _capi.ptr(xyz),
                                              _capi.ptr(P), _capi.ptr(disp), _capi.ptr(mc),
                                              MEM_DEVICE if dev else MEM_HOST))
    return xyz, P, disp, mc


def reproject_poly(disp, coef, R1, min_depth=0.0, max_depth=1000.0, use_rect=1,
                   ctx: Context | None = None):
    """Per-pixel polynomial triangulation of an f64 disparity map."""
    ctx = ctx or default_context()
    dev = _is_dev(disp)
    rows, cols = disp.shape
    device = disp.device if dev else None
    disp = _map("disp", disp, dev, device, torch.float64, (rows, cols))
    xyz = _out(dev, (rows, cols, 3), torch.float32, device)
    P = _out(dev, (rows, cols), torch.uint8, device)
    cx, cy, cz = _coef_maps(coef, dev, rows, cols, device)
    Ra = np.ascontiguousarray(R1, np.float64).reshape(9)
    _capi.check(ctx.lib.psgc_reproject_poly(ctx.h, _capi.ptr(disp), rows, cols, _capi.ptr(cx),
                                            _capi.ptr(cy), _capi.ptr(cz), _capi.ptr(Ra),
                                            float(min_depth), float(max_depth), int(use_rect),
                                            _capi.ptr(xyz), _capi.ptr(P),
                                            MEM_DEVICE if dev else MEM_HOST))
    return xyz, P


def _cloud_in(xyz, valid):
    """An unorganized cloud (N x 3 f32) and its optional mask (N u8)."""
    dev = _is_dev(xyz)
    if dev:
        if xyz.numel() % 3:
            raise ValueError("xyz: expected N x 3 coordinates")
        xyz = _map("xyz", xyz.reshape(-1, 3), True, xyz.device, torch.float32,
                   (xyz.numel() // 3, 3))
    else:
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    if valid is not None:
        n = xyz.shape[0]
        valid = _map("valid", valid.reshape(-1) if dev else np.asarray(valid).reshape(-1), dev,
                     xyz.device if dev else None, torch.uint8, (n,))
    return dev, xyz, valid


def cloud_sor(xyz, valid=None, k=30, std_mul=1.0, voxel=1.0, rmax=4, ctx: Context | None = None):
    """SOR on an unorganized cloud (N x 3 f32): exact k-nearest mean distances
    from a voxel grid; returns (keep u8, mean_dist f64, stats f64[4])."""
    ctx = ctx or default_context()
    dev, xyz, valid = _cloud_in(xyz, valid)
    n = xyz.shape[0]
    device = xyz.device if dev else None
    keep = _out(dev, (n,), torch.uint8, device)
    md = _out(dev, (n,), torch.float64, device)
    st = _out(dev, (4,), torch.float64, device)
    _capi.check(ctx.lib.psgc_cloud_sor(ctx.h, _capi.ptr(xyz), _capi.ptr(valid), n, int(k),
                                       float(std_mul), float(voxel), int(rmax), _capi.ptr(keep),
                                       _capi.ptr(md), _capi.ptr(st), MEM_DEVICE if dev else MEM_HOST))
    return keep, md, st


def cloud_tune_voxel(xyz, valid=None, k=30, candidates=(0.3, 0.4, 0.5, 0.7, 1.0), rmax=4,
                     ctx: Context | None = None):
    """The k-nearest search's voxel autotuner on a device cloud: returns
    (fastest voxel, {voxel: ms})."""
    ctx = (ctx or default_context()).bind_stream()
    dev, xyz, valid = _cloud_in(xyz, valid)
    if not dev:
        raise ValueError("cloud_tune_voxel takes device tensors")
    cand = np.ascontiguousarray(candidates, np.float32)
    times = np.zeros(len(cand), np.float32)
    best = C.c_float()
    _capi.check(ctx.lib.psgc_cloud_tune_voxel(ctx.h, _capi.ptr(xyz), _capi.ptr(valid),
                                              xyz.shape[0], int(k), int(rmax), _capi.ptr(cand),
                                              len(cand), C.byref(best), _capi.ptr(times)))
    # candidates come back as the values passed (not their float32 roundings)
    return (float(candidates[int(np.argmin(np.abs(cand - best.value)))]),
            {float(c): float(t) for c, t in zip(candidates, times)})


def cloud_mls(xyz, valid=None, sigma=0.5, order=2, max_radius=1.0, iterations=1,
              semantics=SEM_GPU, voxel=None, ctx: Context | None = None):
    """MLS projection of an unorganized cloud over the neighbours within
    max_radius (voxel defaults to that radius)."""
    ctx = ctx or default_context()
    dev, xyz, valid = _cloud_in(xyz, valid)
    n = xyz.shape[0]
    r = max_radius if max_radius > 0 else 3.0 * sigma
    out = _out(dev, (n, 3), torch.float32, xyz.device if dev else None)
    _capi.check(ctx.lib.psgc_cloud_mls(ctx.h, _capi.ptr(xyz), _capi.ptr(valid), n, float(sigma),
                                       int(order), float(max_radius), int(iterations),
                                       int(semantics), float(voxel if voxel else r), _capi.ptr(out),
                                       MEM_DEVICE if dev else MEM_HOST))
    return out


def cloud_gaussian(xyz, valid=None, sigma=0.5, radius=1.5, voxel=None,
                   ctx: Context | None = None):
    """Gaussian-weighted mean of each point's neighbours within radius."""
    ctx = ctx or default_context()
    dev, xyz, valid = _cloud_in(xyz, valid)
    n = xyz.shape[0]
    out = _out(dev, (n, 3), torch.float32, xyz.device if dev else None)
    _capi.check(ctx.lib.psgc_cloud_gaussian(ctx.h, _capi.ptr(xyz), _capi.ptr(valid), n,
                                            float(sigma), float(radius),
                                            float(voxel if voxel else radius), _capi.ptr(out),
                                            MEM_DEVICE if dev else MEM_HOST))
    return out


def roi_right_window(lk, rk, pha_f, left_rect, mask, ctx: Context | None = None):
    """prepareRectifiedROI's right search window for left_rect = (x, y, w, h)
    and its mask; returns right_rect as a 4-element int32 array."""
    ctx = ctx or default_context()
    dev = _is_dev(lk)
    rows, cols = lk.shape
    device = lk.device if dev else None
    lr = np.asarray(left_rect, np.int32).reshape(4)
    lk = _map("l_k", lk, dev, device, torch.int32, (rows, cols))
    rk = _map("r_k", rk, dev, device, torch.int32, (rows, cols))
    mask = _map("mask", mask, dev, device, torch.uint8, (int(lr[3]), int(lr[2])))
    rr = np.zeros(4, np.int32)
    _capi.check(ctx.lib.psgc_roi_right_window(ctx.h, _capi.ptr(lk), _capi.ptr(rk), rows, cols,
                                              int(pha_f), _capi.ptr(lr), _capi.ptr(mask),
                                              _capi.ptr(rr), MEM_DEVICE if dev else MEM_HOST))
    return rr


def reconstruct_poly_roi(cfg: ReconConfig, lp, rp, lk, rk, pha_f, left_rect, right_rect, mask,
                         coef, R1, ctx: Context | None = None):
    """startCalculate_v2_roi: reconstruct only the ROI (left_rect, right_rect
    = (x, y, w, h), mask of left_rect's size). Returns crop-sized
    (xyz, P, disp)."""
    ctx = ctx or default_context()
    dev, device, rows, cols, (lp, rp, lk, rk) = _recon_maps(lp, rp, lk, rk)
    lr = np.asarray(left_rect, np.int32).reshape(4)
    rr = np.asarray(right_rect, np.int32).reshape(4)
    h, w = int(lr[3]), int(lr[2])
    mask = _map("mask", mask, dev, device, torch.uint8, (h, w))
    xyz = _out(dev, (h, w, 3), torch.float32, device)
    P = _out(dev, (h, w), torch.uint8, device)
    disp = _out(dev, (h, w), torch.float32, device)
    cx, cy, cz = _coef_maps(coef, dev, rows, cols, device)
    Ra = np.ascontiguousarray(R1, np.float64).reshape(9)
    _capi.check(ctx.lib.psgc_reconstruct_poly_roi(
        ctx.h, C.byref(cfg), _capi.ptr(lp), _capi.ptr(rp), _capi.ptr(lk), _capi.ptr(rk), rows,
        cols, int(pha_f), _capi.ptr(lr), _capi.ptr(rr), _capi.ptr(mask), _capi.ptr(cx),
        _capi.ptr(cy), _capi.ptr(cz), _capi.ptr(Ra), _capi.ptr(xyz), _capi.ptr(P),
        _capi.ptr(disp), MEM_DEVICE if dev else MEM_HOST))
    return xyz, P, disp


def load_coef_bin(path):
    """One _rx/_ry/_rz coefficient file (int32 type, rows, cols + CV_32FC4
    data, JSLAB_StereoCalibParams_V2.cpp:4-17) as a rows x cols x 4 float32
    array."""
    lib = _capi.load()
    r, c = C.c_int(), C.c_int()
    _capi.check(lib.psgc_load_coef_bin(str(path).encode(), C.byref(r), C.byref(c), None))
    out = np.empty((r.value, c.value, 4), np.float32)
    _capi.check(lib.psgc_load_coef_bin(str(path).encode(), C.byref(r), C.byref(c),
                                       _capi.ptr(out)))
    return out


class Batch:
    """Batched decode + reconstruct of many stereo pairs (psgc_batch_*).

    pairs: list of dicts with keys left, right (lists of device frame
    tensors), decode (DecodeConfig), recon (ReconConfig), pha_f, Q, R1 and
    optionally disp=True and coef=(cx, cy, cz) device coefficient maps (the
    polynomial triangulation instead of Q). Outputs (xyz, P[, disp]) are
    allocated here and overwritten by every run()."""

    def __init__(self, pairs, ctx: Context | None = None, graph=True, autotune=False):
        self.ctx = ctx or default_context()
        self.ctx.bind_stream()
        n = len(pairs)
        descs = (_capi.PairDesc * n)()
        self.outputs = []
        self._keep = []
        for i, p in enumerate(pairs):
            L, R = p["left"], p["right"]
            L, ldev, lst, lstride, rows, cols = _frames(L, f"pairs[{i}].left")
            R, rdev, rst, rstride, rrows, rcols = _frames(R, f"pairs[{i}].right")
            if not (ldev and rdev) or R[0].device != L[0].device:
                raise ValueError(f"pairs[{i}]: frames must be device tensors on one device")
            if len(L) != len(R) or len(L) > len(descs[i].left):
                raise ValueError(f"pairs[{i}]: {len(L)} left and {len(R)} right frames "
                                 f"(equal counts of at most {len(descs[i].left)})")
            if (rst, rstride, rrows, rcols) != (lst, lstride, rows, cols):
                raise ValueError(f"pairs[{i}]: right frames differ from the left in type, "
                                 f"shape or row stride")
            d = descs[i]
            for j, f in enumerate(L):
                d.left[j] = f.data_ptr()
            for j, f in enumerate(R):
                d.right[j] = f.data_ptr()
            d.n_frames = len(L)
            d.sample_type = lst
            d.row_stride_bytes = lstride
            d.rows, d.cols = rows, cols
            d.decode = p["decode"]
            d.recon = p["recon"]
            d.pha_f = int(p["pha_f"])
            d.Q[:] = [float(v) for v in np.asarray(p["Q"]).reshape(16)]
            d.R1[:] = [float(v) for v in np.asarray(p["R1"]).reshape(9)]
            dev = L[0].device
            out = {"xyz": torch.empty((rows, cols, 3), dtype=torch.float32, device=dev),
                   "P": torch.empty((rows, cols), dtype=torch.uint8, device=dev)}
            if p.get("disp"):
                out["disp"] = torch.empty((rows, cols), dtype=torch.float32, device=dev)
            d.xyz = out["xyz"].data_ptr()
            d.P = out["P"].data_ptr()
            d.disp = out["disp"].data_ptr() if "disp" in out else None
            if p.get("coef") is not None:
                cm = _coef_maps(p["coef"], True, rows, cols, dev)
                for a in range(3):
                    d.coef[a] = cm[a].data_ptr()
                self._keep.append(cm)
            self.outputs.append(out)
            self._keep.append((L, R))
        flags = (_capi.BATCH_GRAPH if graph else 0) | (_capi.BATCH_AUTOTUNE if autotune else 0)
        h = C.c_void_p()
        _capi.check(self.ctx.lib.psgc_batch_create(self.ctx.h, descs, n, flags, C.byref(h)))
        self.h = h

    def run(self):
        self.ctx.bind_stream()
        _capi.check(self.ctx.lib.psgc_batch_run(self.h))
        return self.outputs

    def run_eager(self):
        self.ctx.bind_stream()
        _capi.check(self.ctx.lib.psgc_batch_run_eager(self.h))
        return self.outputs

    def grid(self, stage: str) -> int:
        g = C.c_int()
        _capi.check(self.ctx.lib.psgc_batch_grid(self.h, stage.encode(), C.byref(g)))
        return g.value

    def set_grid(self, stage: str, value: int):
        """Pins a stage's launch shape (re-captures the graph)."""
        self.ctx.bind_stream()
        _capi.check(self.ctx.lib.psgc_batch_set_grid(self.h, stage.encode(), int(value)))

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.psgc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------ reference-shaped classes --

class Version(enum.Enum):
    V1 = 0
    V2 = 1


class PhaseShiftDecoder:
    """JSLAB_PhaseShiftDecoder (JSLAB_PhaseShiftDecoder.h:10-92) on the GPU."""

    Version = Version

    def __init__(self, config: DecodeConfig, ctx: Context | None = None):
        self.config = config
        self.ctx = ctx
        self._images = None
        self._k = self._mod = self._valid = None

    @staticmethod
    def create(config: DecodeConfig, version: Version = Version.V1):
        """Returns None when any frequency is zero, as the reference factory
        does (JSLAB_PhaseShiftDecoder.cpp:9-29). V1 has no order map and is
        superseded; both versions map to the V2 decoder here."""
        if config.unwrap_mode == UNWRAP_HETERODYNE3 and any(f == 0 for f in config.freq):
            return None
        return PhaseShiftDecoder(config)

    def setImages(self, images):
        n = sum(self.config.shift_num[:3]) if self.config.unwrap_mode == UNWRAP_HETERODYNE3 \
            else self.config.shift_num[0] + 2 * self.config.gray_bits
        if len(images) != n:
            raise ValueError(f"expected {n} images, got {len(images)}")
        self._images = list(images)

    def startCalculate(self):
        ph, self._k, self._mod, self._valid = decode(self.config, self._images, self.ctx)
        return ph

    def getOrderMap(self):
        return self._k

    def getModulationMap(self):
        return self._mod

    def getValidMask(self):
        return self._valid


def stereo_rectify(K1, K2, R, T, image_size):
    """cv::stereoRectify(K1, 0, K2, 0, size, R, T, flags 0, alpha -1) as
    setCalibrationParams calls it (JSLAB_PhaseEpipolarReconstructor_V2.cpp:
    302-307), zero distortion (psgc_stereo_rectify): returns R1, R2 (3x3),
    P1, P2 (3x4) and Q (4x4)."""
    lib = _capi.load()
    K1, K2, R = (np.ascontiguousarray(a, np.float64).reshape(9) for a in (K1, K2, R))
    T = np.ascontiguousarray(T, np.float64).reshape(3)
    R1, R2, Q = np.zeros(9), np.zeros(9), np.zeros(16)
    P1, P2 = np.zeros(12), np.zeros(12)
    _capi.check(lib.psgc_stereo_rectify(_capi.ptr(K1), _capi.ptr(K2), _capi.ptr(R), _capi.ptr(T),
                                        int(image_size[0]), int(image_size[1]), _capi.ptr(R1),
                                        _capi.ptr(R2), _capi.ptr(P1), _capi.ptr(P2), _capi.ptr(Q)))
    return R1.reshape(3, 3), R2.reshape(3, 3), P1.reshape(3, 4), P2.reshape(3, 4), Q.reshape(4, 4)


@dataclasses.dataclass
class StereoCalibParams_V1:
    """JSLAB_StereoCalibParams_V1's fields (camera matrices, distortion,
    R, t of the right camera, image size (width, height)); only zero
    distortion is supported."""
    camera_matrix: tuple
    dist_coeffs: tuple
    R: object
    t: object
    image_size: tuple


def full_left_rectification(K1, image_size, R1, P1, P2, Q):
    """computeFullLeftRectification (JSLAB_PhaseEpipolarReconstructor_V2.cpp:
    248-289) for zero distortion: returns (P1, P2, Q, (width, height)) with
    the principal points shifted so that the whole left image lands on the
    rectified canvas, and the canvas size (psgc_full_left_rectification)."""
    lib = _capi.load()
    K1 = np.ascontiguousarray(K1, np.float64).reshape(9)
    R1 = np.ascontiguousarray(R1, np.float64).reshape(9)
    P1 = np.array(P1, np.float64).reshape(12)
    P2 = np.array(P2, np.float64).reshape(12)
    Q = np.array(Q, np.float64).reshape(16)
    w, h = C.c_int(), C.c_int()
    _capi.check(lib.psgc_full_left_rectification(_capi.ptr(K1), int(image_size[0]),
                                                 int(image_size[1]), _capi.ptr(R1), _capi.ptr(P1),
                                                 _capi.ptr(P2), _capi.ptr(Q), C.byref(w),
                                                 C.byref(h)))
    return P1.reshape(3, 4), P2.reshape(3, 4), Q.reshape(4, 4), (w.value, h.value)


def rectify_maps(K, R, P, size, ctx: Context | None = None, device="cuda"):
    """cv::initUndistortRectifyMap(K, 0, R, P, size, CV_32FC1) on the device
    (psgc_rectify_maps): two f32 maps (height, width) of source pixel
    coordinates."""
    ctx = ctx or default_context()
    w, h = int(size[0]), int(size[1])
    K = np.ascontiguousarray(K, np.float64).reshape(9)
    R = np.ascontiguousarray(R, np.float64).reshape(9)
    Kr = np.ascontiguousarray(np.asarray(P, np.float64).reshape(3, -1)[:, :3]).reshape(9)
    mx = torch.empty((h, w), dtype=torch.float32, device=device)
    my = torch.empty_like(mx)
    _capi.check(ctx.lib.psgc_rectify_maps(ctx.h, _capi.ptr(K), _capi.ptr(R), _capi.ptr(Kr), h, w,
                                          _capi.ptr(mx), _capi.ptr(my), MEM_DEVICE))
    return mx, my


def remap(src, map_x, map_y, nearest=False, ctx: Context | None = None):
    """cv::remap as the reference's setData calls it (…V2.cpp:372-376): a f64
    map with INTER_LINEAR, or an i32 map with INTER_NEAREST (nearest=True),
    BORDER_CONSTANT 0 (psgc_remap). Device tensors in, a device tensor of the
    maps' shape out."""
    ctx = ctx or default_context()
    if not (_is_dev(src) and _is_dev(map_x) and _is_dev(map_y)):
        raise ValueError("remap: src and maps must be device tensors")
    rows, cols = map_x.shape
    dev = src.device
    dt = torch.int32 if nearest else torch.float64
    if src.dim() != 2:
        raise ValueError(f"remap: src must be 2-D, got {tuple(src.shape)}")
    src = _map("src", src, True, dev, dt, tuple(src.shape))
    map_x = _map("map_x", map_x, True, dev, torch.float32, (rows, cols))
    map_y = _map("map_y", map_y, True, dev, torch.float32, (rows, cols))
    dst = torch.empty((rows, cols), dtype=dt, device=dev)
    _capi.check(ctx.lib.psgc_remap(ctx.h, 1 if nearest else 0, _capi.ptr(src), src.shape[0],
                                   src.shape[1], _capi.ptr(map_x), _capi.ptr(map_y), rows, cols,
                                   _capi.ptr(dst), MEM_DEVICE))
    return dst


class PhaseEpipolarReconstructor_V2:
    """JSLAB_PhaseEpipolarReconstructor_V2 (JSLAB_PhaseEpipolarReconstructor_V2.h:
    11-171) on the GPU. The calibration is given either as Q and R1 for a
    pre-rectified pair (setCalibrationParams; setData then takes the maps as
    they are), or as the stereoRectify outputs of a zero-distortion pair
    (setRectification: the full-left canvas when the config asks for it, and
    device remap tables that setData applies as the reference does)."""

    def __init__(self, config: ReconConfig, ctx: Context | None = None):
        self.config = config
        self.ctx = ctx
        self.Q = self.R1 = None
        self._disp = None
        self._maps = None
        self.output_size = None

    def setCalibrationParams(self, params, R1=None, device="cuda"):
        """setCalibrationParams (…V2.cpp:292-338) from a StereoCalibParams_V1
        with zero distortion: stereoRectify, the full-left canvas when the
        config asks for it, then the remap tables. setCalibrationParams(Q, R1)
        instead takes a pre-rectified pair's Q and R1 (setData then uses the
        maps as they are)."""
        if R1 is not None:
            self.Q = np.asarray(params, np.float64).reshape(16)
            self.R1 = np.asarray(R1, np.float64).reshape(9)
            self._maps = None
            return
        p = params
        if any(np.any(np.asarray(d, np.float64) != 0) for d in p.dist_coeffs):
            raise NotImplementedError("setCalibrationParams: only zero distortion is supported")
        K1, K2 = (np.asarray(k, np.float64) for k in p.camera_matrix)
        R1_, R2_, P1, P2, Q = stereo_rectify(K1, K2, p.R, p.t, p.image_size)
        self.setRectification(K1, K2, R1_, R2_, P1, P2, Q, p.image_size, device)

    def setRectification(self, K1, K2, R1, R2, P1, P2, Q, image_size, device="cuda"):
        """The part of setCalibrationParams after cv::stereoRectify
        (…V2.cpp:313-338) for zero distortion: with use_full_left_rectification
        the principal points move and the canvas grows to hold the whole left
        image (computeFullLeftRectification), then the left and right remap
        tables are built on the device (initUndistortRectifyMap)."""
        size = (int(image_size[0]), int(image_size[1]))
        if self.config.use_full_left_rectification:
            P1, P2, Q, size = full_left_rectification(K1, image_size, R1, P1, P2, Q)
        self.Q = np.asarray(Q, np.float64).reshape(16)
        self.R1 = np.asarray(R1, np.float64).reshape(9)
        self.output_size = size
        self._maps = (rectify_maps(K1, R1, P1, size, self.ctx, device),
                      rectify_maps(K2, R2, P2, size, self.ctx, device))

    def setData(self, left, right, left_k, right_k, left_mod, pha_f):
        """setData (…V2.cpp:368-395): phase and modulation remapped with
        INTER_LINEAR, fringe orders with INTER_NEAREST, when remap tables are
        set."""
        if self._maps is not None:
            (lx, ly), (rx, ry) = self._maps
            left, right = remap(left, lx, ly, ctx=self.ctx), remap(right, rx, ry, ctx=self.ctx)
            left_k = remap(left_k, lx, ly, nearest=True, ctx=self.ctx)
            right_k = remap(right_k, rx, ry, nearest=True, ctx=self.ctx)
            if left_mod is not None:
                left_mod = remap(left_mod, lx, ly, ctx=self.ctx)
        self._in = (left, right, left_k, right_k, int(pha_f))
        self._mod = left_mod

    def startCalculate(self):
        lp, rp, lk, rk, pha_f = self._in
        xyz, P, self._disp, self._match = reconstruct(self.config, lp, rp, lk, rk, pha_f, self.Q,
                                                      self.R1, self.ctx)
        return xyz, P

    def getDisparityMap(self):
        return self._disp


def applyMLSFilterToPointCloud(src_xyz, P, window_size, spatial_sigma, use_gpu=True,
                               polynomial_order=0, max_radius=0.0, iterations=1,
                               semantics=SEM_GPU, ctx: Context | None = None):
    """applyMLSFilterToPointCloud (JSLAB_MLS_Filter.h:53-57); the GPU path
    always runs (there is no CPU implementation in this package)."""
    ctx = ctx or default_context()
    dev = _is_dev(src_xyz)
    rows, cols = P.shape
    device = src_xyz.device if dev else None
    src_xyz = _map("src_xyz", src_xyz, dev, device, torch.float32, (rows, cols, 3))
    P = _map("P", P, dev, device, torch.uint8, (rows, cols))
    dst = _out(dev, (rows, cols, 3), torch.float32, device)
    _capi.check(ctx.lib.psgc_filter_mls(ctx.h, _capi.ptr(src_xyz), _capi.ptr(P), _capi.ptr(dst),
                                        rows, cols, int(window_size), float(spatial_sigma),
                                        int(polynomial_order), float(max_radius),
                                        int(iterations), int(semantics),
                                        MEM_DEVICE if dev else MEM_HOST))
    return dst
