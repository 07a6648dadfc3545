"""setData's cv::remap (JSLAB_PhaseEpipolarReconstructor_V2.cpp:368-377) and
the full-left rectification canvas (…V2.cpp:248-289), on the CPU.

The oracle's remap restates OpenCV 4.7.0's rule for CV_32FC1 maps (the
reference's pinned OpenCV, absent here): coordinates rounded half to even
to 1/32 px for INTER_LINEAR, to 1 px for INTER_NEAREST, BORDER_CONSTANT 0.
It is pinned by what that rule fixes in closed form: on an integer ramp the
output is the ramp at the quantised coordinate, exactly; on an image of
ones near the border it is the inside weight; ties go to even. A plausible
mistake (exact coordinates, truncation, ties away from zero, swapped
weights) fails one of them, which the planted-mistake test shows.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from zim_amd import _capi

INT_MIN = np.iinfo(np.int32).min


def q32(v):
    """cvRound(v * 32) as OpenCV converts a float map: half to even, INT_MIN
    for NaN, infinities and values outside int range."""
    s = np.asarray(v, np.float32) * np.float32(32)
    ok = (s >= -2147483648.0) & (s < 2147483648.0)
    return np.where(ok, np.rint(np.where(ok, s, 0)).astype(np.int64), INT_MIN)


def numpy_remap_linear(src, mx, my):
    """The same rule written independently with numpy (vectorised, in the
    table's order)."""
    X, Y = q32(mx), q32(my)
    sx, sy = np.clip(X >> 5, -32768, 32767), np.clip(Y >> 5, -32768, 32767)
    ax = (X & 31).astype(np.float32) / np.float32(32)
    ay = (Y & 31).astype(np.float32) / np.float32(32)
    one = np.float32(1)
    w = [(one - ay) * (one - ax), (one - ay) * ax, ay * (one - ax), ay * ax]
    rows, cols = src.shape

    def tap(y, x):
        inside = (x >= 0) & (x < cols) & (y >= 0) & (y < rows)
        return np.where(inside, src[np.clip(y, 0, rows - 1), np.clip(x, 0, cols - 1)], 0.0)

    v = tap(sy, sx) * w[0].astype(np.float64)
    v = v + tap(sy, sx + 1) * w[1].astype(np.float64)
    v = v + tap(sy + 1, sx) * w[2].astype(np.float64)
    return v + tap(sy + 1, sx + 1) * w[3].astype(np.float64)


def special_maps(rows, cols, seed):
    g = np.random.default_rng(seed)
    mx = g.uniform(-3, cols + 2, (rows, cols)).astype(np.float32)
    my = g.uniform(-3, rows + 2, (rows, cols)).astype(np.float32)
    # ties of the 1/32 grid and of the integer grid, both signs
    mx[0, :8] = np.array([1 / 64, 3 / 64, 5 / 64, -1 / 64, 0.5, 1.5, -0.5, 2.5], np.float32)
    my[1, :6] = np.array([1 / 64, 3 / 64, -3 / 64, 0.5, -1.5, 1e-9], np.float32)
    mx[2, :6] = np.array([np.nan, np.inf, -np.inf, 3e9, -3e9, 1e6], np.float32)
    my[3, :4] = np.array([np.nan, -np.inf, 5e9, -40000.0], np.float32)
    return mx, my


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_remap_equals_numpy_rule(seed):
    rows, cols = 23, 31
    g = np.random.default_rng(100 + seed)
    src = g.normal(0, 10, (17, 29))
    mx, my = special_maps(rows, cols, seed)
    got = oracle.remap(src, mx, my)
    want = numpy_remap_linear(src, mx, my)
    assert np.array_equal(got, want)
    assert np.all(got[2, :3] == 0) and np.all(got[3, :3] == 0)


def test_linear_on_ramp_is_ramp_at_quantised_coordinate():
    rows, cols = 40, 50
    yy, xx = np.mgrid[0:rows, 0:cols]
    a, b, c = 7.0, -3.0, 11.0
    src = a * xx + b * yy + c
    g = np.random.default_rng(5)
    mx = g.uniform(0, cols - 1, (30, 30)).astype(np.float32)
    my = g.uniform(0, rows - 1, (30, 30)).astype(np.float32)
    mx[0, :4] = [1 / 64, 3 / 64, 47.5 / 32, 0.0]
    my[0, :4] = [5 / 64, 1 / 64, 0.0, 31 / 32 + 1 / 64]
    got = oracle.remap(src, mx, my)
    qx, qy = q32(mx) / 32.0, q32(my) / 32.0
    assert np.array_equal(got, a * qx + b * qy + c)
    # ties went to even
    assert np.array_equal(q32(np.float32([1 / 64, 3 / 64, 5 / 64, -1 / 64, -3 / 64])),
                          [0, 2, 2, 0, -2])
    # and the quantisation is visible: up to 1/64 px away from the exact ramp
    dev = np.abs(got - (a * mx + b * my + c))
    assert dev.max() > 0.05 and dev.max() <= (abs(a) + abs(b)) / 64 + 1e-9


def test_linear_border_reads_zero_outside():
    src = np.ones((6, 7))
    xs = np.float32([-1.25, -1.0, -0.75, -0.25, -1 / 32, 0.0, 5.5, 6.0, 6.25, 6.75, 7.0, 7.5])
    mx = np.tile(xs, (1, 1))
    my = np.full_like(mx, 2.0)
    got = oracle.remap(src, mx, my)[0]
    q = q32(xs) / 32.0
    # inside weight: 1 + q on [-1, 0], 1 inside, (cols - q) - 0 on [cols-1, cols]
    want = np.where(q <= -1, 0.0, np.where(q < 0, 1 + q, np.where(q <= 6, 1.0,
                                                                  np.maximum(7 - q, 0.0))))
    assert np.array_equal(got, want)


def test_nearest_rounds_half_to_even():
    src = np.arange(1, 1 + 5 * 6, dtype=np.int32).reshape(5, 6)
    mx = np.float32([[0.5, 1.5, 2.5, -0.5, -0.51, 5.49, 5.5, np.nan, 2.0, 2.0]])
    my = np.float32([[0.0, 0.0, 0.0, 0.0, 0.0, 4.5, 0.0, 0.0, -0.5, 4.51]])
    got = oracle.remap(src, mx, my, nearest=True)[0]
    want = [src[0, 0], src[0, 2], src[0, 2], src[0, 0], 0, src[4, 5], 0, 0, src[0, 2], 0]
    assert list(got) == want


def test_planted_mistakes_fail_the_ramp():
    """Variants of the rule that a restatement could plausibly get wrong,
    each caught by the closed forms above."""
    rows, cols = 20, 20
    yy, xx = np.mgrid[0:rows, 0:cols]
    src = 7.0 * xx - 3.0 * yy
    g = np.random.default_rng(9)
    mx = g.uniform(1, cols - 2, (16, 16)).astype(np.float32)
    my = g.uniform(1, rows - 2, (16, 16)).astype(np.float32)
    mx[0, 0], my[0, 0] = 5 / 64, 0.0  # a tie: 2.5 / 32 goes to 2 / 32, not 3 / 32
    want = 7.0 * (q32(mx) / 32.0) - 3.0 * (q32(my) / 32.0)
    assert np.array_equal(oracle.remap(src, mx, my), want)
    exact = 7.0 * mx.astype(np.float64) - 3.0 * my.astype(np.float64)  # no quantisation
    trunc = 7.0 * np.floor(mx * 32.0) / 32.0 - 3.0 * np.floor(my * 32.0) / 32.0
    away = 7.0 * np.floor(mx * 32.0 + 0.5) / 32.0 - 3.0 * np.floor(my * 32.0 + 0.5) / 32.0
    swapped = -3.0 * (q32(mx) / 32.0) + 7.0 * (q32(my) / 32.0)  # transposed weights
    for wrong in (exact, trunc, away, swapped):
        assert not np.array_equal(wrong, want)


# ---------------------------------------------------------- full-left canvas --

def _full_left(K1, size, R1, P1, P2, Q):
    lib = _capi.load()
    P1, P2, Q = (np.array(a, np.float64).reshape(-1) for a in (P1, P2, Q))
    w, h = C.c_int(), C.c_int()
    rc = lib.psgc_full_left_rectification(_capi.ptr(np.ascontiguousarray(K1, np.float64)),
                                          size[0], size[1],
                                          _capi.ptr(np.ascontiguousarray(R1, np.float64)),
                                          _capi.ptr(P1), _capi.ptr(P2), _capi.ptr(Q),
                                          C.byref(w), C.byref(h))
    return rc, P1.reshape(3, 4), P2.reshape(3, 4), Q.reshape(4, 4), (w.value, h.value)


def _stereo(K1, f, cx, cy, B):
    P1 = np.array([[f, 0, cx, 0], [0, f, cy, 0], [0, 0, 1, 0]], np.float64)
    P2 = P1.copy()
    P2[0, 3] = -f * B
    Q = np.array([[1, 0, 0, -cx], [0, 1, 0, -cy], [0, 0, 0, f], [0, 0, 1 / B, 0]], np.float64)
    return P1, P2, Q


def _roty(deg):
    t = np.radians(deg)
    return np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])


def test_full_left_identity_keeps_the_canvas():
    # dyadic intrinsics: (0 - cx) / f * f + cx is exactly 0, so the corners
    # map onto themselves and nothing moves
    K1 = np.array([[2048.0, 0, 1224.0], [0, 2048.0, 1024.0], [0, 0, 1]])
    P1, P2, Q = _stereo(K1, 2048.0, 1224.0, 1024.0, 150.0)
    rc, P1n, P2n, Qn, size = _full_left(K1, (2448, 2048), np.eye(3), P1, P2, Q)
    assert rc == 0 and size == (2448, 2048)
    assert np.array_equal(P1n, P1) and np.array_equal(P2n, P2) and np.array_equal(Qn, Q)
    # otherwise the (0, 0) corner comes back within rounding of 0, on either
    # side, and a value just below 0 moves the canvas by one pixel (the
    # reference's arithmetic does the same)
    K1 = np.array([[2400.0, 0, 1223.5], [0, 2400.0, 1023.5], [0, 0, 1]])
    P1, P2, Q = _stereo(K1, 2400.0, 1223.5, 1023.5, 150.0)
    rc, P1n, P2n, Qn, size = _full_left(K1, (2448, 2048), np.eye(3), P1, P2, Q)
    s = P1n[:2, 2] - P1[:2, 2]
    assert rc == 0 and size == (2448, 2048) and set(s) <= {0.0, 1.0}
    assert np.array_equal(Qn[:2, 3] - Q[:2, 3], -s)


@pytest.mark.parametrize("deg,f_new", [(2.0, 2300.0), (-3.0, 2500.0), (0.7, 2400.0)])
def test_full_left_canvas_spans_every_corner(deg, f_new):
    W, H = 1280, 1024
    K1 = np.array([[2400.0, 0, 630.0], [0, 2390.0, 515.0], [0, 0, 1]])
    R1 = _roty(deg) @ np.array([[1, 0, 0], [0, np.cos(0.01), -np.sin(0.01)],
                                [0, np.sin(0.01), np.cos(0.01)]])
    P1, P2, Q = _stereo(K1, f_new, 640.0, 512.0, 80.0)
    rc, P1n, P2n, Qn, (w, h) = _full_left(K1, (W, H), R1, P1, P2, Q)
    assert rc == 0
    corners = np.array([[0, 0, 1], [W - 1, 0, 1], [W - 1, H - 1, 1], [0, H - 1, 1]], float).T
    img = P1n[:, :3] @ R1 @ np.linalg.inv(K1) @ corners
    u, v = img[0] / img[2], img[1] / img[2]
    # the shifted corners start in [0, 1) and span at most the canvas size:
    # every corner lies in [0, w) x [0, h) (the canvas is the extent, rounded
    # up, from a whole-pixel shift, so up to one pixel of the far edge can sit
    # past the last column), and the canvas is tight
    assert u.min() >= -1e-3 and v.min() >= -1e-3
    assert u.max() < w + 1e-3 and v.max() < h + 1e-3
    assert u.min() < 1 and v.min() < 1
    assert 0 <= w - (u.max() - u.min() + 1) < 1 + 1e-3
    assert 0 <= h - (v.max() - v.min() + 1) < 1 + 1e-3
    # both principal points move by the same whole-pixel shift, so
    # disparities are kept, and Q maps a shifted pixel to the point the
    # unshifted one gave
    s = P1n[:, 2] - P1[:, 2]
    assert np.array_equal(P2n[:, 2] - P2[:, 2], s) and s[2] == 0
    assert np.array_equal(s[:2], np.round(s[:2])) and abs(deg) < 1 or np.any(s != 0)
    for (uu, vv, d) in ((100.0, 200.0, 35.0), (900.0, 10.0, 80.0)):
        a = Q @ np.array([uu, vv, d, 1.0])
        b = Qn @ np.array([uu + s[0], vv + s[1], d, 1.0])
        assert np.allclose(a[:3] / a[3], b[:3] / b[3], rtol=0, atol=1e-9)


def test_full_left_rejects_bad_input():
    K1 = np.eye(3)
    P1, P2, Q = _stereo(K1, 1.0, 0.0, 0.0, 1.0)
    assert _full_left(np.zeros((3, 3)), (10, 10), np.eye(3), P1, P2, Q)[0] == _capi.PSGC_E_INVALID
    assert _full_left(K1, (0, 10), np.eye(3), P1, P2, Q)[0] == _capi.PSGC_E_INVALID


# ------------------------------------------------------ stereo rectification --

def _rodrigues(axis, deg):
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    t = np.radians(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K


def _rig(seed, vertical=False):
    g = np.random.default_rng(seed)
    K1 = np.array([[2400 + g.uniform(-30, 30), 0, 1224 + g.uniform(-20, 20)],
                   [0, 2400 + g.uniform(-30, 30), 1024 + g.uniform(-20, 20)], [0, 0, 1]])
    K2 = np.array([[2400 + g.uniform(-30, 30), 0, 1224 + g.uniform(-20, 20)],
                   [0, 2400 + g.uniform(-30, 30), 1024 + g.uniform(-20, 20)], [0, 0, 1]])
    R = _rodrigues(g.normal(size=3), g.uniform(1, 6))
    T = np.array([-150.0, g.uniform(-8, 8), g.uniform(-8, 8)])
    if vertical:
        T = T[[1, 0, 2]]
    return K1, K2, R, T


@pytest.mark.parametrize("seed,vertical", [(1, False), (2, False), (3, False), (4, True)])
def test_stereo_rectify_geometry(seed, vertical):
    """What stereoRectify must achieve, from the geometry alone: both
    rectified cameras share one orientation (R2 R R1^T = I), the baseline
    lies on the x axis (y for a vertical rig), rows (columns) of a scene point
    agree between the two rectified images, and Q takes a rectified pixel and
    its disparity back to the point."""
    import zim_amd
    K1, K2, R, T = _rig(seed, vertical)
    R1, R2, P1, P2, Q = zim_amd.stereo_rectify(K1, K2, R, T, (2448, 2048))
    for Rk in (R1, R2):
        assert np.abs(Rk @ Rk.T - np.eye(3)).max() < 1e-12 and abs(np.linalg.det(Rk) - 1) < 1e-12
    assert np.abs(R2 @ R @ R1.T - np.eye(3)).max() < 1e-12
    t = R2 @ T
    ax = 1 if vertical else 0
    assert np.abs(np.delete(t, ax)).max() < 1e-9 * np.linalg.norm(T)
    assert P1[0, 0] == P1[1, 1] == P2[0, 0] == P2[1, 1]
    assert P2[ax, 3] == pytest.approx(t[ax] * P1[0, 0], rel=1e-12)
    assert P1[1 - ax, 2] == P2[1 - ax, 2]
    g = np.random.default_rng(seed + 50)
    X = np.column_stack([g.uniform(-200, 200, 40), g.uniform(-150, 150, 40), g.uniform(400, 900, 40)])
    # raw pixels of scene points (left camera frame), then their rectified pixels
    ul = (K1 @ X.T).T
    ur = (K2 @ (R @ X.T + T[:, None])).T
    rl = (P1[:, :3] @ R1 @ np.linalg.inv(K1) @ ul.T).T
    rr = (P2[:, :3] @ R2 @ np.linalg.inv(K2) @ ur.T).T
    rl, rr = rl[:, :2] / rl[:, 2:], rr[:, :2] / rr[:, 2:]
    assert np.abs(rl[:, 1 - ax] - rr[:, 1 - ax]).max() < 1e-6
    d = rl[:, ax] - rr[:, ax]
    h = Q @ np.column_stack([rl, d, np.ones(len(d))]).T
    Xq = (h[:3] / h[3]).T
    assert np.abs(Xq - (R1 @ X.T).T).max() < 1e-6 * 900


def test_stereo_rectify_of_an_ideal_rig_is_the_identity():
    """An already rectified rig (R = I, T along x, equal focal lengths)
    keeps its cameras and principal points, and Q is the closed form."""
    import zim_amd
    K = np.array([[2400.0, 0, 1223.5], [0, 2400.0, 1023.5], [0, 0, 1]])
    R1, R2, P1, P2, Q = zim_amd.stereo_rectify(K, K, np.eye(3), [-150.0, 0, 0], (2448, 2048))
    assert np.abs(R1 - np.eye(3)).max() < 1e-15 and np.abs(R2 - np.eye(3)).max() < 1e-15
    assert np.abs(P1[:, :3] - K).max() < 1e-3
    lib = _capi.load()
    Qi, R1i = np.zeros(16), np.zeros(9)
    assert lib.psgc_rectify_ideal(2400.0, 2400.0, P1[0, 2], P1[1, 2], P2[0, 2], 150.0,
                                  _capi.ptr(Qi), _capi.ptr(R1i)) == 0
    assert np.abs(Q.ravel() - Qi).max() < 1e-12


def test_stereo_rectify_half_turn():
    """Cameras turned by a half turn about a skew axis: the half rotation
    each camera takes still aligns them (the rotation vector's axis and sign
    are recovered at theta = pi)."""
    import zim_amd
    K = np.array([[2400.0, 0, 1223.5], [0, 2400.0, 1023.5], [0, 0, 1]])
    for axis in ([0.2, 1.0, 0.1], [0.0, 0.6, 0.8], [1.0, -0.3, 0.2]):
        R = _rodrigues(axis, 180.0)
        R1, R2, P1, P2, Q = zim_amd.stereo_rectify(K, K, R, [-150.0, 3.0, 1.0], (2448, 2048))
        assert np.isfinite(Q).all()
        # the angle comes from acos of the trace, whose slope is unbounded at
        # pi: about sqrt(eps) of angle error (a wrong axis or sign is O(1))
        assert np.abs(R2 @ R @ R1.T - np.eye(3)).max() < 1e-6, axis


def test_calibration_params_object_rejects_distortion():
    import zim_amd
    K = np.eye(3)
    p = zim_amd.StereoCalibParams_V1((K, K), (np.zeros(5), np.array([0.1, 0, 0, 0, 0])),
                                     np.eye(3), np.array([-1.0, 0, 0]), (64, 48))
    with pytest.raises(NotImplementedError):
        zim_amd.PhaseEpipolarReconstructor_V2(_capi.ReconConfig.default()).setCalibrationParams(p)
