"""Small seeded scenes shared by the tests (rendered on the CPU)."""
from __future__ import annotations

import functools

import numpy as np

from zim_amd import synth
from zim_amd._capi import DecodeConfig, ReconConfig, SEM_GPU, OUTLIER_SOR

C2_SMALL = dict(rows=192, cols=256)


@functools.lru_cache(maxsize=None)
def c2_small(seed=3, rows=192, cols=256):
    """Config-2 shaped pair (u16 12-bit, 70/64/59 heterodyne, 4 steps) cut
    down so the CPU restatement runs in well under a second: the full-size
    focal length (so the 0.2-0.4 mm pixel pitch and the filter radii keep
    their meaning) with a short 20 mm baseline (so the disparity fits the
    narrow image)."""
    p = synth.render_c2(seed, rows=rows, cols=cols, f=2400.0, baseline=20.0)
    left = [x.numpy() for x in p.left]
    right = [x.numpy() for x in p.right]
    return p, left, right


@functools.lru_cache(maxsize=None)
def c1_small(seed=5, gray=True):
    p = synth.render_c1(seed, rows=120, cols=160, with_gray=gray)
    return p, [x.numpy() for x in p.left], [x.numpy() for x in p.right]


def c2_decode_config(**kw):
    base = dict(freq=(70, 64, 59), shift_num=(4, 4, 4), modulation_threshold=0.3)
    base.update(kw)
    return DecodeConfig.default(**base)


def c1_gray_config(**kw):
    base = dict(freq=(50.0, 0.0, 0.0), shift_num=(4, 0, 0), modulation_threshold=0.3,
                unwrap_mode=1, gray_bits=6, gray_period_px=16.0, gray_min_contrast=20.0)
    base.update(kw)
    return DecodeConfig.default(**base)


def c2_recon_config(**kw):
    base = dict(use_mls=1, mls_polynomial_order=2, mls_max_radius=1.5, mls_spatial_sigma=0.5,
                mls_window_size=6, mls_iterations=2, semantics=SEM_GPU, gauss_window=5,
                gauss_sigma=1.0, outlier_mode=OUTLIER_SOR, sor_k=30, sor_std_mul=1.0,
                sor_window=4)
    base.update(kw)
    return ReconConfig.default_gpu(**base)


def rng_points(rows, cols, seed=0, holes=0.1, z0=700.0):
    """A smooth organized cloud with holes, for filter-stage tests."""
    g = np.random.default_rng(seed)
    ii, jj = np.mgrid[0:rows, 0:cols].astype(np.float64)
    z = z0 + 20 * np.sin(ii / 17.0) * np.cos(jj / 23.0) + g.normal(0, 0.05, (rows, cols))
    x = (jj - cols / 2) * z / 2400.0
    y = (ii - rows / 2) * z / 2400.0
    xyz = np.stack([x, y, z], -1).astype(np.float32)
    P = (g.random((rows, cols)) > holes).astype(np.uint8)
    # a few far outliers
    out = g.random((rows, cols)) < 0.01
    xyz[out, 2] += g.normal(0, 5, out.sum()).astype(np.float32)
    xyz[P == 0] = 0
    return xyz, P


HDR_ROWS = ((120.0, 100.0), (150.0, 140.0), (400.0, 100.0))  # (A, B) per row block


def hdr_sweep(cols=2000, block=3, steps=8, freqs=(70.0, 64.0, 59.0), width=1280.0, sat=255.0):
    """Noise-free f64 three-frequency fringes I_k = A + B cos(phi_f + 2 pi k / N)
    over a sweep of projector columns x in [0, width), phi_f = 2 pi f x / width,
    in three blocks of `block` rows with the (A, B) of HDR_ROWS: unclipped,
    partly clipped (at most 2 of 8 samples reach `sat`) and saturated (every
    sample reaches `sat`). Samples at or above `sat` are set to `sat`.
    Returns (frames, phi0 per column, fringe order per column)."""
    x = (np.arange(cols, dtype=np.float64) + 0.37) * width / cols
    frames = []
    for f in freqs:
        phi = 2 * np.pi * f * x / width
        for k in range(steps):
            img = np.empty((3 * block, cols))
            for b, (A, B) in enumerate(HDR_ROWS):
                img[b * block:(b + 1) * block] = A + B * np.cos(phi + 2 * np.pi * k / steps)
            frames.append(np.minimum(img, sat))
    phi0 = 2 * np.pi * freqs[0] * x / width
    order = np.round(phi0 / (2 * np.pi)).astype(np.int32)
    return frames, phi0, order


def hdr_config(**kw):
    base = dict(freq=(70, 64, 59), shift_num=(8, 8, 8), modulation_threshold=0.3, enable_hdr=1,
                hdr_saturation_gray=255.0, hdr_min_amplitude=2.0)
    base.update(kw)
    return DecodeConfig.default(**base)


@functools.lru_cache(maxsize=None)
def hdr_small(seed=6, rows=192, cols=256):
    """An over-exposed 8-step config-2 shaped pair in uint8: the 12-bit
    intensities scaled by 255/1500 and clipped, so that bright fringe samples
    saturate at 255 (the case the HDR fit exists for)."""
    p = synth.render_c2(seed, rows=rows, cols=cols, f=2400.0, baseline=20.0, steps=(8, 8, 8))

    def to_u8(t):
        v = np.round(t.numpy().astype(np.float64) * (255.0 / 1500.0))
        return np.clip(v, 0, 255).astype(np.uint8)

    return p, [to_u8(x) for x in p.left], [to_u8(x) for x in p.right]


def q_ideal(f, cx1, cy, cx2, B):
    """Q of an ideal rectified pair (cv::stereoRectify's layout): w = (d -
    (cx1 - cx2)) / B, X = (j - cx1)/w, Y = (i - cy)/w, Z = f/w."""
    return np.array([[1, 0, 0, -cx1], [0, 1, 0, -cy], [0, 0, 0, f],
                     [0, 0, 1.0 / B, -(cx1 - cx2) / B]], np.float64)


def poly_coefs_equivalent_to_q(rows, cols, f, cx1, cy, cx2, B, form):
    """Coefficient maps whose polynomial triangulation is the ideal Q
    reprojection, in one of three exact forms:
      'inv_cubic'  cx.w = -1: 1/z = (d - D)/(f B), x = (j - cx1)/f z, y = (i - cy)/f z
      'rational'   cx.w = -3, cz.w = -4: z = (-f B / D) / (1 - d / D)
      'inv_quad'   cz.w = -1: 1/z as above (quadratic term 0)
    with D = cx1 - cx2 (non-zero for 'rational')."""
    ii, jj = np.mgrid[0:rows, 0:cols].astype(np.float64)
    D = cx1 - cx2
    cx = np.zeros((rows, cols, 4), np.float32)
    cy_ = np.zeros((rows, cols, 4), np.float32)
    cz = np.zeros((rows, cols, 4), np.float32)
    cx[..., 0] = (jj - cx1) / f
    cy_[..., 0] = (ii - cy) / f
    if form == "inv_cubic":
        cx[..., 3] = -1.0
        cz[..., 0] = -D / (f * B)
        cz[..., 1] = 1.0 / (f * B)
    elif form == "rational":
        cx[..., 3] = -3.0
        cz[..., 0] = -f * B / D
        cz[..., 1] = -1.0 / D
        cz[..., 3] = -4.0
    elif form == "inv_quad":
        cz[..., 0] = -D / (f * B)
        cz[..., 1] = 1.0 / (f * B)
        cz[..., 3] = -1.0
    else:
        raise ValueError(form)
    return cx, cy_, cz


@functools.lru_cache(maxsize=None)
def closeup(seed=41, rows=128, cols=192, noise=4.0):
    """A close-range pair at ~0.025 mm pixel pitch (a gently curved surface
    near Z = 60 mm; f = 2400 px, B = 2 mm, 80 px of disparity): the scale of
    the reference GPU class's defaults (ROR radius 0.05 mm with 5 of an 11 x 11
    window, MLS at 3 sigma = 1.5 mm in a 51 x 51 window,
    JSLAB_PhaseEpipolarReconstructor_V2_GPU.h:27-61). 12-bit frames of the
    config-2 patterns with N(0, noise DN)."""
    import torch
    cam = synth.Camera(rows, cols, 2400.0, 2.0)
    proj = synth.Projector(1280, 1000.0)
    ii = torch.arange(rows, dtype=torch.float64).view(-1, 1)
    jj = torch.arange(cols, dtype=torch.float64).view(1, -1)
    z = 60.0 + 0.004 * (jj - cam.cx) + 0.3 * torch.sin(jj / 23.0) * torch.cos(ii / 17.0)
    p = synth._render(cam, proj, z, seed, "cpu", (70, 64, 59), (4, 4, 4), ambient=400.0,
                      modulation=750.0, noise=noise, poisson=False, max_val=4095,
                      dtype=torch.uint16)
    return p, [x.numpy() for x in p.left], [x.numpy() for x in p.right]


def ring_points(xyz, P, hw, radius):
    """Valid points with a valid neighbour within `radius` on the outer ring
    of their (2 hw + 1)^2 window (Chebyshev distance exactly hw): 0 means a
    window one pixel wider would select no further neighbour, i.e. the
    window holds the radius for this cloud. numpy arrays or torch tensors
    (on any device); distances in FP64."""
    import torch
    if isinstance(xyz, torch.Tensor):
        X = xyz.double()
        valid = P != 0
        hit = torch.zeros_like(valid)
    else:
        X = np.asarray(xyz, np.float64)
        valid = np.asarray(P) > 0
        hit = np.zeros_like(valid)
    rows, cols = valid.shape
    r2 = float(radius) ** 2
    for dy in range(-hw, hw + 1):
        for dx in range(-hw, hw + 1):
            if max(abs(dy), abs(dx)) != hw:
                continue
            ys, ye = max(0, -dy), min(rows, rows - dy)
            xs, xe = max(0, -dx), min(cols, cols - dx)
            a = X[ys:ye, xs:xe]
            b = X[ys + dy:ye + dy, xs + dx:xe + dx]
            both = valid[ys:ye, xs:xe] & valid[ys + dy:ye + dy, xs + dx:xe + dx]
            hit[ys:ye, xs:xe] |= both & (((a - b) ** 2).sum(-1) <= r2)
    return int(hit.sum())


def c2_window(pair, radius=1.5):
    """The MLS half window of a config-2 pair: holds `radius` at the pair's
    nearest depth."""
    import zim_amd
    return zim_amd.mls_half_window(radius, pair.cam.f, float(pair.z_left.min()))
