"""Device-side rectification maps and remap, valid-point compaction."""
import math

import numpy as np
import pytest

import oracle
import torch
import zim_amd
from tests import scenes
from tests.test_remap_cpu import special_maps
from zim_amd import _capi

pytestmark = pytest.mark.gpu


def test_rectify_maps(ctx):
    """initUndistortRectifyMap for zero distortion: the exact homography
    K R^T K_rect^-1, rounded once to float."""
    rows, cols = 60, 80
    K = np.array([[700.0, 0, 39.5], [0, 700.0, 29.5], [0, 0, 1]])
    th = math.radians(0.5)
    R = np.array([[math.cos(th), 0, math.sin(th)], [0, 1, 0], [-math.sin(th), 0, math.cos(th)]])
    Krect = np.array([[690.0, 0, 41.0], [0, 690.0, 28.0], [0, 0, 1]])
    mx = np.zeros((rows, cols), np.float32)
    my = np.zeros((rows, cols), np.float32)
    _capi.check(ctx.lib.psgc_rectify_maps(ctx.h, _capi.ptr(K), _capi.ptr(R), _capi.ptr(Krect),
                                          rows, cols, _capi.ptr(mx), _capi.ptr(my),
                                          _capi.MEM_HOST))
    uu, vv = np.meshgrid(np.arange(cols), np.arange(rows))
    ray = np.linalg.inv(Krect) @ np.stack([uu.ravel(), vv.ravel(), np.ones(uu.size)])
    raw = K @ (R.T @ ray)
    ex, ey = (raw[0] / raw[2]).reshape(rows, cols), (raw[1] / raw[2]).reshape(rows, cols)
    # within one float rounding (half an ulp, plus the FP64 evaluation)
    for got, want in ((mx, ex), (my, ey)):
        ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(got - want) <= ulp * 0.5 + 1e-9)


@pytest.mark.parametrize("seed", [0, 1])
def test_remap_matches_oracle(ctx, seed):
    """cv::remap's rule (OpenCV 4.7.0: 1/32 px INTER_LINEAR, INTER_NEAREST,
    BORDER_CONSTANT 0) on the device equals the oracle bit for bit, on
    ties, NaN, infinite and far-out coordinates included."""
    g = np.random.default_rng(300 + seed)
    src = g.normal(0, 10, (37, 45))
    ks = g.integers(0, 70, (37, 45)).astype(np.int32)
    mx, my = special_maps(41, 53, seed)
    dst = np.zeros(mx.shape)
    _capi.check(ctx.lib.psgc_remap(ctx.h, 0, _capi.ptr(src), 37, 45, _capi.ptr(mx),
                                   _capi.ptr(my), 41, 53, _capi.ptr(dst), _capi.MEM_HOST))
    assert np.array_equal(dst, oracle.remap(src, mx, my))
    kd = np.zeros(mx.shape, np.int32)
    _capi.check(ctx.lib.psgc_remap(ctx.h, 1, _capi.ptr(ks), 37, 45, _capi.ptr(mx),
                                   _capi.ptr(my), 41, 53, _capi.ptr(kd), _capi.MEM_HOST))
    assert np.array_equal(kd, oracle.remap(ks, mx, my, nearest=True))
    # device tensors through the Python API
    t = zim_amd.remap(torch.from_numpy(src).cuda(), torch.from_numpy(mx).cuda(),
                      torch.from_numpy(my).cuda(), ctx=ctx)
    assert np.array_equal(t.cpu().numpy(), dst)


def _rot(ax, deg):
    t = math.radians(deg)
    c, s = math.cos(t), math.sin(t)
    if ax == "x":
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


@pytest.mark.parametrize("full_left", [1, 0])
def test_reconstructor_with_rectification_matches_oracle(ctx, full_left):
    """The reference's setCalibrationParams + setData + startCalculate with
    a rotated zero-distortion rectification: full-left canvas, device remap
    tables, then the chain; against the oracle's remap and chain on the
    same tables."""
    p, left, right = scenes.c2_small(3)
    dcfg = scenes.c2_decode_config()
    L, R = oracle.decode(dcfg, left), oracle.decode(dcfg, right)
    rows, cols = L[0].shape
    f = 2400.0
    K1 = np.array([[f, 0, cols / 2 - 0.5], [0, f, rows / 2 - 0.5], [0, 0, 1]])
    K2 = K1.copy()
    R1 = _rot("y", 0.2) @ _rot("x", 0.1)
    R2 = R1.copy()
    Q = np.asarray(p.Q, np.float64).reshape(4, 4)
    P1 = np.array([[f, 0, cols / 2 + 3.25, 0], [0, f, rows / 2 - 2.75, 0], [0, 0, 1, 0]])
    P2 = P1.copy()
    cfg = scenes.c2_recon_config(use_full_left_rectification=full_left, mls_window_size=4)
    rec = zim_amd.PhaseEpipolarReconstructor_V2(cfg, ctx)
    rec.setRectification(K1, K2, R1, R2, P1, P2, Q, (cols, rows))
    w, h = rec.output_size
    if full_left:
        assert (w, h) != (cols, rows)
    rec.setData(*(torch.from_numpy(np.ascontiguousarray(a)).cuda()
                  for a in (L[0], R[0], L[1], R[1], L[2])), 70)
    gx, gP = rec.startCalculate()
    torch.cuda.synchronize()
    (lx, ly), (rx, ry) = [(a.cpu().numpy(), b.cpu().numpy()) for a, b in rec._maps]
    lp, rp = oracle.remap(L[0], lx, ly), oracle.remap(R[0], rx, ry)
    lk, rk = oracle.remap(L[1], lx, ly, nearest=True), oracle.remap(R[1], rx, ry, nearest=True)
    assert np.array_equal(rec._in[0].cpu().numpy(), lp)
    assert np.array_equal(rec._in[3].cpu().numpy(), rk)
    ox, oP, od, omc = oracle.reconstruct(cfg, lp, rp, lk, rk, 70, rec.Q, rec.R1)
    gx, gP = gx.cpu().numpy(), gP.cpu().numpy()
    assert gP.shape == (h, w)
    assert np.array_equal(gP, oP), f"mask mismatch {(gP != oP).sum()}"
    assert np.abs(gx - ox).max() <= 1e-3
    assert oP.mean() > 0.2


@pytest.mark.parametrize("shape,density", [((37, 53), 0.6), ((512, 700), 0.3), ((1, 1), 1.0),
                                           ((5, 5), 0.0)])
def test_compact_stable_order(ctx, shape, density):
    g = np.random.default_rng(shape[0])
    xyz = g.normal(0, 100, shape + (3,)).astype(np.float32)
    P = (g.random(shape) < density).astype(np.uint8)
    gray = g.integers(0, 256, shape).astype(np.uint8)
    n = int(P.sum())
    out = np.zeros((max(P.size, 1), 3), np.float32)
    og = np.zeros(max(P.size, 1), np.uint8)
    import ctypes as C
    cnt = C.c_int64()
    _capi.check(ctx.lib.psgc_compact(ctx.h, _capi.ptr(xyz), _capi.ptr(P), _capi.ptr(gray),
                                     shape[0], shape[1], _capi.ptr(out), _capi.ptr(og),
                                     C.byref(cnt), _capi.MEM_HOST))
    assert cnt.value == n
    assert np.array_equal(out[:n], xyz[P > 0])
    assert np.array_equal(og[:n], gray[P > 0])


def test_reconstructor_from_calibration_params_matches_oracle(ctx):
    """setCalibrationParams from the stereo calibration (K1, K2, R, t, zero
    distortion): stereoRectify, the full-left canvas and the remap tables on
    the device; setData and startCalculate against the oracle's remap and
    chain on the same tables and matrices."""
    p, left, right = scenes.c2_small(5)
    dcfg = scenes.c2_decode_config()
    L, R = oracle.decode(dcfg, left), oracle.decode(dcfg, right)
    rows, cols = L[0].shape
    K = np.array([[2400.0, 0, cols / 2 - 0.5], [0, 2400.0, rows / 2 - 0.5], [0, 0, 1]])
    Rrig = _rot("y", 0.15) @ _rot("x", -0.1)
    params = zim_amd.StereoCalibParams_V1((K, K), (np.zeros(5), np.zeros(5)), Rrig,
                                          np.array([-20.0, 0.1, 0.05]), (cols, rows))
    cfg = scenes.c2_recon_config(use_full_left_rectification=1, mls_window_size=4)
    rec = zim_amd.PhaseEpipolarReconstructor_V2(cfg, ctx)
    rec.setCalibrationParams(params)
    R1, R2, P1, P2, Q = zim_amd.stereo_rectify(K, K, Rrig, params.t, (cols, rows))
    assert np.allclose(rec.R1, R1.ravel())
    rec.setData(*(torch.from_numpy(np.ascontiguousarray(a)).cuda()
                  for a in (L[0], R[0], L[1], R[1], L[2])), 70)
    gx, gP = rec.startCalculate()
    torch.cuda.synchronize()
    (lx, ly), (rx, ry) = [(a.cpu().numpy(), b.cpu().numpy()) for a, b in rec._maps]
    lp, rp = oracle.remap(L[0], lx, ly), oracle.remap(R[0], rx, ry)
    lk, rk = oracle.remap(L[1], lx, ly, nearest=True), oracle.remap(R[1], rx, ry, nearest=True)
    ox, oP, _, _ = oracle.reconstruct(cfg, lp, rp, lk, rk, 70, rec.Q, rec.R1)
    gx, gP = gx.cpu().numpy(), gP.cpu().numpy()
    assert gP.shape == lx.shape
    assert np.array_equal(gP, oP), f"mask mismatch {(gP != oP).sum()}"
    assert np.abs(gx - ox).max() <= 1e-3
    assert oP.sum() > 1000
