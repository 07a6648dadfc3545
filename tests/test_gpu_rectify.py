"""Device-side rectification maps and remap, valid-point compaction."""
import math

import numpy as np
import pytest

from zim_amd import _capi

pytestmark = pytest.mark.gpu


def test_rectify_maps_and_remap(ctx):
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
    assert np.abs(mx - ex).max() < 1e-3 and np.abs(my - ey).max() < 1e-3
    g = np.random.default_rng(0)
    src = g.normal(0, 1, (rows, cols))
    dst = np.zeros_like(src)
    _capi.check(ctx.lib.psgc_remap(ctx.h, 0, _capi.ptr(src), rows, cols, _capi.ptr(mx),
                                   _capi.ptr(my), rows, cols, _capi.ptr(dst), _capi.MEM_HOST))
    x0 = np.floor(mx.astype(np.float64)).astype(int)
    y0 = np.floor(my.astype(np.float64)).astype(int)
    a = mx.astype(np.float64) - x0
    b = my.astype(np.float64) - y0
    pad = np.pad(src, 2)
    tap = lambda yy, xx: np.where((xx >= 0) & (xx < cols) & (yy >= 0) & (yy < rows),  # noqa: E731
                                  pad[np.clip(yy, -2, rows + 1) + 2, np.clip(xx, -2, cols + 1) + 2], 0)
    ref = ((1 - a) * (1 - b) * tap(y0, x0) + a * (1 - b) * tap(y0, x0 + 1)
           + (1 - a) * b * tap(y0 + 1, x0) + a * b * tap(y0 + 1, x0 + 1))
    assert np.abs(dst - ref).max() < 1e-12
    ks = g.integers(0, 70, (rows, cols)).astype(np.int32)
    kd = np.zeros_like(ks)
    _capi.check(ctx.lib.psgc_remap(ctx.h, 1, _capi.ptr(ks), rows, cols, _capi.ptr(mx),
                                   _capi.ptr(my), rows, cols, _capi.ptr(kd), _capi.MEM_HOST))
    xi, yi = np.rint(mx).astype(int), np.rint(my).astype(int)
    inside = (xi >= 0) & (xi < cols) & (yi >= 0) & (yi < rows)
    refk = np.where(inside, ks[np.clip(yi, 0, rows - 1), np.clip(xi, 0, cols - 1)], 0)
    assert np.array_equal(kd, refk)


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
