"""Seeded synthetic rectified stereo pairs with projected fringe patterns.

The reference ships no data (its paths are absolute Windows directories), so
every workload here is rendered. One renderer serves the CPU tests (torch on
the CPU, small sizes) and the benchmark (torch on the GPU, 5 MP pairs).

Geometry (all lengths in mm, left camera frame, rectified pair):
  left camera at the origin, right camera at (B, 0, 0), both with focal f
  (pixels) and principal point (cx, cy); disparity d = x_L - x_R = f B / Z.
  The projector is a pinhole at (B/2, 0, 0) with the cameras' orientation,
  focal fp and principal column cp; a surface point (X, Y, Z) sees projector
  column u_p = fp (X - B/2) / Z + cp, and fringe level i has phase
  2 pi freq[i] u_p / proj_width. Vertical fringes, so the phase runs along the
  epipolar rows.
The scene is a depth map Z(i, j) over the left image; the right view is
rasterised from it row by row with a z-buffer (nearest surface wins), so
occlusions are handled. Projector shadows are not modelled.

Intensities follow I = A + M (1 + cos(phi + 2 pi k / N)) with ambient A and
modulation M (the literal I = A + M cos of a workload description would clip
below zero), then noise and clipping to the sample range.
"""
from __future__ import annotations

import dataclasses
import math

import torch


@dataclasses.dataclass
class Camera:
    rows: int
    cols: int
    f: float
    baseline: float
    cx: float | None = None
    cy: float | None = None

    def __post_init__(self):
        if self.cx is None:
            self.cx = (self.cols - 1) / 2.0
        if self.cy is None:
            self.cy = (self.rows - 1) / 2.0


@dataclasses.dataclass
class Projector:
    width: int
    f: float
    cp: float | None = None

    def __post_init__(self):
        if self.cp is None:
            self.cp = (self.width - 1) / 2.0


def q_matrix(cam: Camera) -> list[float]:
    """Q of cv::stereoRectify for this ideal pair (cx1 == cx2), row-major."""
    return [1.0, 0.0, 0.0, -cam.cx,
            0.0, 1.0, 0.0, -cam.cy,
            0.0, 0.0, 0.0, cam.f,
            0.0, 0.0, 1.0 / cam.baseline, 0.0]


IDENTITY3 = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]


# ---------------------------------------------------------------- scenes ---

def depth_bumps(cam: Camera, seed: int, device="cpu", n_bumps=64,
                z_range=(500.0, 900.0)):
    """Plane at Z0 ~ U(z_range) plus n_bumps Gaussian bumps (sigma 20-120 px,
    height +-(5..50) mm) defined in left-image pixel coordinates."""
    g = torch.Generator().manual_seed(seed)
    z0 = z_range[0] + (z_range[1] - z_range[0]) * torch.rand(1, generator=g, dtype=torch.float64).item()
    tilt = (torch.rand(2, generator=g, dtype=torch.float64) - 0.5) * 0.1  # mm per px
    cy_ = torch.rand(n_bumps, generator=g, dtype=torch.float64) * cam.rows
    cx_ = torch.rand(n_bumps, generator=g, dtype=torch.float64) * cam.cols
    sg = 20.0 + 100.0 * torch.rand(n_bumps, generator=g, dtype=torch.float64)
    amp = (5.0 + 45.0 * torch.rand(n_bumps, generator=g, dtype=torch.float64))
    amp = amp * torch.where(torch.rand(n_bumps, generator=g) < 0.5, -1.0, 1.0).double()
    ii = torch.arange(cam.rows, dtype=torch.float64, device=device).view(-1, 1)
    jj = torch.arange(cam.cols, dtype=torch.float64, device=device).view(1, -1)
    z = z0 + tilt[0].item() * (jj - cam.cx) + tilt[1].item() * (ii - cam.cy)
    z = z.expand(cam.rows, cam.cols).clone()
    for m in range(n_bumps):
        r2 = (ii - cy_[m].item()) ** 2 + (jj - cx_[m].item()) ** 2
        z -= amp[m].item() * torch.exp(-r2 / (2.0 * sg[m].item() ** 2))
    return z


def depth_plane_spheres(cam: Camera, seed: int, device="cpu", n_spheres=3,
                        z_plane=600.0, tilt_deg=20.0):
    """Plane tilted tilt_deg about the vertical axis through (0, 0, z_plane)
    plus n_spheres spheres of radius U(30, 80) mm, ray-cast from the left
    camera (nearest hit)."""
    g = torch.Generator().manual_seed(seed)
    ii = torch.arange(cam.rows, dtype=torch.float64, device=device).view(-1, 1)
    jj = torch.arange(cam.cols, dtype=torch.float64, device=device).view(1, -1)
    dx = ((jj - cam.cx) / cam.f).expand(cam.rows, cam.cols)
    dy = ((ii - cam.cy) / cam.f).expand(cam.rows, cam.cols)
    # plane: n . P = n . P0 with n = (sin t, 0, cos t); ray P = Z (dx, dy, 1)
    t = math.radians(tilt_deg)
    n = (math.sin(t), 0.0, math.cos(t))
    z = (n[2] * z_plane) / (n[0] * dx + n[2])
    for _ in range(n_spheres):
        r = 30.0 + 50.0 * torch.rand(1, generator=g, dtype=torch.float64).item()
        zc = z_plane - 60.0 - 120.0 * torch.rand(1, generator=g, dtype=torch.float64).item()
        half_w = (cam.cols / 2.0) / cam.f * zc * 0.7
        half_h = (cam.rows / 2.0) / cam.f * zc * 0.7
        xc = (torch.rand(1, generator=g, dtype=torch.float64).item() * 2 - 1) * half_w
        yc = (torch.rand(1, generator=g, dtype=torch.float64).item() * 2 - 1) * half_h
        # |Z d - C|^2 = r^2, d = (dx, dy, 1)
        a = dx * dx + dy * dy + 1.0
        b = -2.0 * (dx * xc + dy * yc + zc)
        c = xc * xc + yc * yc + zc * zc - r * r
        disc = b * b - 4 * a * c
        hit = disc >= 0
        zs = (-b - torch.sqrt(torch.clamp(disc, min=0.0))) / (2 * a)
        z = torch.where(hit & (zs > 0) & (zs < z), zs, z)
    return z


# ------------------------------------------------------------- rendering ---

def projector_column(cam: Camera, proj: Projector, jstar, z):
    """u_p of the surface point seen at left column jstar with depth z."""
    return proj.f * ((jstar - cam.cx) / cam.f - cam.baseline / (2.0 * z)) + proj.cp


def right_view(cam: Camera, proj: Projector, z_left):
    """Rasterises the right view of the left depth map. Returns (u_p, disp)
    per right pixel (NaN where no surface is seen)."""
    rows, cols = z_left.shape
    dev = z_left.device
    fb = cam.f * cam.baseline
    jj = torch.arange(cols, dtype=torch.float64, device=dev).view(1, -1)
    xr = jj - fb / z_left                       # right column of each left sample
    x0, x1 = xr[:, :-1], xr[:, 1:]
    z0, z1 = z_left[:, :-1], z_left[:, 1:]
    seg_ok = torch.isfinite(z0) & torch.isfinite(z1) & (x1 > x0)
    span = torch.where(seg_ok, x1 - x0, torch.zeros_like(x0))
    m_max = int(min(16, math.ceil(float(span.max().item())) + 1)) if seg_ok.any() else 1
    best = torch.full((rows * cols,), -math.inf, dtype=torch.float64, device=dev)
    row_base = (torch.arange(rows, device=dev) * cols).view(-1, 1)
    cands = []
    for m in range(m_max):
        r = torch.ceil(x0) + m
        ok = seg_ok & (r <= x1) & (r >= 0) & (r < cols)
        t = (r - x0) / torch.where(seg_ok, x1 - x0, torch.ones_like(x0))
        zs = z0 + t * (z1 - z0)
        disp = fb / zs
        idx = (row_base + r.clamp(0, cols - 1).long())
        sel = ok.nonzero(as_tuple=True)
        flat_idx = idx[sel]
        d_sel = disp[sel]
        best.scatter_reduce_(0, flat_idx, d_sel, reduce="amax", include_self=True)
        up = projector_column(cam, proj, (jj[:, :-1] + t)[sel], zs[sel])
        cands.append((flat_idx, d_sel, up))
    up_r = torch.full((rows * cols,), math.nan, dtype=torch.float64, device=dev)
    for flat_idx, d_sel, up in cands:
        win = d_sel == best[flat_idx]
        up_r[flat_idx[win]] = up[win]
    disp_r = torch.where(torch.isfinite(best), best, torch.full_like(best, math.nan))
    return up_r.view(rows, cols), disp_r.view(rows, cols)


def fringe_frames(up, proj: Projector, freqs, steps, ambient, modulation,
                  gen, noise_sigma, poisson, max_val, dtype, lowmod=None):
    """Phase-shift frames for each level: I = A + M (1 + cos(phi + 2 pi k/N))."""
    lit = torch.isfinite(up) & (up >= 0) & (up < proj.width)
    mod = torch.where(lit, torch.full_like(up, modulation), torch.zeros_like(up))
    if lowmod is not None:
        mod = torch.where(lowmod, mod * 0.05, mod)
    upz = torch.where(lit, up, torch.zeros_like(up))
    frames = []
    for f, n in zip(freqs, steps):
        phi = 2.0 * math.pi * f * upz / proj.width
        for k in range(n):
            frames.append(ambient + mod * (1.0 + torch.cos(phi + 2.0 * math.pi * k / n)))
    return [_noisy(I, gen, noise_sigma, poisson, max_val, dtype) for I in frames]


def gray_frames(up, proj: Projector, period, bits, dark, bright, gen,
                noise_sigma, max_val, dtype):
    """Gray-code stripe frames (MSB first) and their complements; stripe
    n = floor(u_p / period) carries code n ^ (n >> 1)."""
    lit = torch.isfinite(up) & (up >= 0) & (up < proj.width)
    n = torch.floor(torch.where(lit, up, torch.zeros_like(up)) / period).long()
    g = n ^ (n >> 1)
    pos, neg = [], []
    for b in range(bits):
        bit = ((g >> (bits - 1 - b)) & 1).double()
        on = torch.where(lit, dark + (bright - dark) * bit, torch.full_like(up, dark))
        off = torch.where(lit, dark + (bright - dark) * (1 - bit), torch.full_like(up, dark))
        pos.append(_noisy(on, gen, noise_sigma, False, max_val, dtype))
        neg.append(_noisy(off, gen, noise_sigma, False, max_val, dtype))
    return pos + neg


def _noisy(I, gen, sigma, poisson, max_val, dtype):
    if poisson:
        I = torch.poisson(I, generator=gen)
    if sigma > 0:
        I = I + sigma * torch.randn(I.shape, generator=gen, dtype=I.dtype, device=I.device)
    return torch.clamp(torch.round(I), 0, max_val).to(dtype)


@dataclasses.dataclass
class Pair:
    cam: Camera
    proj: Projector
    left: list            # frames (torch tensors, rows x cols)
    right: list
    z_left: torch.Tensor  # f64 truth depth per left pixel
    freqs: tuple
    steps: tuple
    Q: list
    R1: list


def render_c2(seed: int, device="cpu", rows=2048, cols=2448, f=2400.0,
              baseline=150.0, freqs=(70, 64, 59), steps=(4, 4, 4),
              noise_seed: int | None = None) -> Pair:
    """Full-size single-pair workload: 2448x2048, bump height field on a plane
    at 500-900 mm, three-frequency heterodyne 4-step fringes (70/64/59 across
    a 1280-px projector), uint16 12-bit, ambient 400, modulation 750,
    Poisson + N(0, 4 DN) noise. noise_seed (default: seed) seeds the noise
    alone, so several views of one surface differ only in their noise."""
    cam = Camera(rows, cols, f, baseline)
    proj = Projector(1280, 1000.0 * f / 2400.0)  # projector/camera magnification of 1000/2400
    z = depth_bumps(cam, seed, device)
    return _render(cam, proj, z, seed if noise_seed is None else noise_seed, device, freqs,
                   steps, ambient=400.0, modulation=750.0, noise=4.0, poisson=True,
                   max_val=4095, dtype=torch.uint16)


@dataclasses.dataclass
class Scene:
    """World frame in mm, z up; the plane z = 0 carries every object."""
    spheres: list   # (cx, cy, cz, r)
    boxes: list     # (cx, cy, cz, hx, hy, hz, yaw)


def c5_scene(seed: int, n_objects=12, extent=180.0) -> Scene:
    """One plane plus n_objects resting on it, alternately spheres (radius
    20-50 mm) and boxes (half extents 15-40 mm, yaw 0-90 deg), centres
    uniform over +-extent mm in x and y, seeded."""
    g = torch.Generator().manual_seed(seed)
    u = lambda: torch.rand(1, generator=g, dtype=torch.float64).item()  # noqa: E731
    spheres, boxes = [], []
    for i in range(n_objects):
        x, y = (2 * u() - 1) * extent, (2 * u() - 1) * extent
        if i % 2 == 0:
            r = 20.0 + 30.0 * u()
            spheres.append((x, y, r, r))
        else:
            hx, hy, hz = 15.0 + 25.0 * u(), 15.0 + 25.0 * u(), 15.0 + 25.0 * u()
            boxes.append((x, y, hz, hx, hy, hz, u() * math.pi / 2))
    return Scene(spheres, boxes)


def c5_pose(seed: int, view: int, n_views=16):
    """Left-camera pose of one view: on a ring around the scene (azimuth
    360 view / n_views deg, +-5 deg), elevation 55-75 deg, 650-800 mm from a
    target within 30 mm of the origin, looking at it (OpenCV axes: x right, y
    down, z forward). Returns (R, C): R maps world directions to the camera
    frame, C is the camera centre (FP64 tensors)."""
    g = torch.Generator().manual_seed(seed * 1000 + view)
    u = lambda: torch.rand(1, generator=g, dtype=torch.float64).item()  # noqa: E731
    az = 2 * math.pi * view / n_views + math.radians(10.0 * u() - 5.0)
    el = math.radians(55.0 + 20.0 * u())
    dist = 650.0 + 150.0 * u()
    tgt = torch.tensor([(2 * u() - 1) * 30.0, (2 * u() - 1) * 30.0, 0.0], dtype=torch.float64)
    C = tgt + dist * torch.tensor([math.cos(el) * math.cos(az), math.cos(el) * math.sin(az),
                                   math.sin(el)], dtype=torch.float64)
    fwd = (tgt - C) / torch.linalg.norm(tgt - C)
    right = torch.linalg.cross(fwd, torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64))
    right = right / torch.linalg.norm(right)
    down = torch.linalg.cross(fwd, right)
    return torch.stack([right, down, fwd]), C


def raycast(scene: Scene, R, C, cam: Camera, device="cpu"):
    """Camera-frame depth (Z) of the nearest surface seen by each left pixel
    (inf where none) and the index of what it hit (0 the plane, 1.. the
    spheres then the boxes, -1 nothing)."""
    ii = torch.arange(cam.rows, dtype=torch.float64, device=device).view(-1, 1)
    jj = torch.arange(cam.cols, dtype=torch.float64, device=device).view(1, -1)
    dc = torch.stack(torch.broadcast_tensors((jj - cam.cx) / cam.f, (ii - cam.cy) / cam.f,
                                             torch.ones_like(ii * jj)), -1)
    Rd, Cd = R.to(device), C.to(device)
    D = dc @ Rd       # world direction per pixel (R^T d); its camera Z is 1
    best = torch.full((cam.rows, cam.cols), math.inf, dtype=torch.float64, device=device)
    obj = torch.full((cam.rows, cam.cols), -1, dtype=torch.int32, device=device)

    def take(t, idx):
        nonlocal best, obj
        hit = torch.isfinite(t) & (t > 0) & (t < best)
        best = torch.where(hit, t, best)
        obj = torch.where(hit, torch.full_like(obj, idx), obj)

    take(-Cd[2] / D[..., 2], 0)
    k = 1
    for (x, y, z, r) in scene.spheres:
        o = Cd - torch.tensor([x, y, z], dtype=torch.float64, device=device)
        a = (D * D).sum(-1)
        b = 2 * (D * o).sum(-1)
        c = (o * o).sum() - r * r
        disc = b * b - 4 * a * c
        t = torch.where(disc >= 0, (-b - torch.sqrt(disc.clamp(min=0))) / (2 * a),
                        torch.full_like(a, math.inf))
        take(t, k)
        k += 1
    for (x, y, z, hx, hy, hz, yaw) in scene.boxes:
        cy_, sy_ = math.cos(yaw), math.sin(yaw)
        Rb = torch.tensor([[cy_, sy_, 0.0], [-sy_, cy_, 0.0], [0.0, 0.0, 1.0]],
                          dtype=torch.float64, device=device)   # world -> box
        o = Rb @ (Cd - torch.tensor([x, y, z], dtype=torch.float64, device=device))
        d = D @ Rb.T
        h = torch.tensor([hx, hy, hz], dtype=torch.float64, device=device)
        inv = 1.0 / d
        t1, t2 = (-h - o) * inv, (h - o) * inv
        tmin = torch.minimum(t1, t2).amax(-1)
        tmax = torch.maximum(t1, t2).amin(-1)
        t = torch.where((tmax >= tmin) & (tmin > 0), tmin, torch.full_like(tmin, math.inf))
        take(t, k)
        k += 1
    return best, obj


def render_c5_view(scene_seed: int, view: int, device="cpu", rows=2048, cols=2448,
                   f=2400.0, baseline=150.0, n_views=16):
    """One view of the multi-object scene (c5_scene) from c5_pose: the
    rectified pair rendered as the full-size workload renders its own (70/64/59
    4-step fringes on a 1280-px projector between the cameras, uint16 12-bit,
    ambient 400, modulation 750, Poisson + N(0, 4 DN)). Returns the pair,
    the pose (R, C) and the object index per left pixel."""
    cam = Camera(rows, cols, f, baseline)
    proj = Projector(1280, 1000.0 * f / 2400.0)
    R, C = c5_pose(scene_seed, view, n_views)
    z, obj = raycast(c5_scene(scene_seed), R, C, cam, device)
    pair = _render(cam, proj, z, scene_seed * 100 + view, device, (70, 64, 59), (4, 4, 4),
                   ambient=400.0, modulation=750.0, noise=4.0, poisson=True, max_val=4095,
                   dtype=torch.uint16)
    return pair, R, C, obj


def to_world(xyz, R, C):
    """Camera-frame points (..., 3) of a view into the scene's world frame:
    R^T x + C (FP64 pose, FP32 points out)."""
    R = R.to(xyz.device)
    Cd = C.to(xyz.device)
    return (xyz.double() @ R + Cd).float()


def render_ref_timing(seed: int, device="cpu", size=5120) -> Pair:
    """The configuration of the reference's published timing
    (JSLAB_Dual_GPU移植说明.md:128-133): a size x size 8-bit pair, three
    frequencies 107/100/94 with 8 steps each (24 frames per camera). Scene:
    the C2 bump height field at the same angular pixel size (f scaled with
    the width), ambient 25, modulation 45 (I in [25, 115]), N(0, 2 DN)."""
    f = 2400.0 * size / 2448.0
    cam = Camera(size, size, f, 150.0)
    proj = Projector(1280, 1000.0 * f / 2400.0)
    z = depth_bumps(cam, seed, device)
    return _render(cam, proj, z, seed, device, (107, 100, 94), (8, 8, 8), ambient=25.0,
                   modulation=45.0, noise=2.0, poisson=False, max_val=255, dtype=torch.uint8)


def render_c1(seed: int, device="cpu", rows=480, cols=640, with_gray=True) -> Pair:
    """CPU-runnable workload: 640x480, f = 800 px, B = 100 mm, plane tilted
    20 deg at 600 mm plus 3 spheres, 4-step fringes of period 16 px on an
    800-px projector, 6-bit Gray code with complements, uint8, ambient 50,
    modulation 100, N(0, 2 DN) noise, 2 % low-modulation pixels."""
    cam = Camera(rows, cols, 800.0, 100.0)
    proj = Projector(800, 700.0 * cols / 640.0)
    z = depth_plane_spheres(cam, seed, device)
    pair = _render(cam, proj, z, seed, device, (800 / 16.0,), (4,), ambient=50.0,
                   modulation=100.0, noise=2.0, poisson=False, max_val=255,
                   dtype=torch.uint8, lowmod_frac=0.02, gray=(16.0, 6) if with_gray else None)
    return pair


def render_c3(seed: int, device="cpu", rows=480, cols=640, steps=4, gray=False) -> Pair:
    """Uneven-batch pair: the config-1 scene (plane tilted 20 deg at 600 mm
    plus 3 spheres, f = 800 px per 640 columns, B = 100 mm) at any
    resolution, uint16 12-bit (ambient 400, modulation 750, Poisson +
    N(0, 4 DN), 2 % low-modulation pixels), with `steps`-step fringes:
    three-frequency 70/64/59 across a 1280-px projector, or (gray) period
    16 px on an 800-px projector plus the 6-bit Gray code with complements.
    Both projectors span the same field (focal 700 px per 800 projector
    columns at 640 camera columns)."""
    f = 800.0 * cols / 640.0
    cam = Camera(rows, cols, f, 100.0)
    z = depth_plane_spheres(cam, seed, device)
    common = dict(ambient=400.0, modulation=750.0, noise=4.0, poisson=True, max_val=4095,
                  dtype=torch.uint16, lowmod_frac=0.02)
    if gray:
        proj = Projector(800, 700.0 * cols / 640.0)
        return _render(cam, proj, z, seed, device, (800 / 16.0,), (steps,), gray=(16.0, 6),
                       **common)
    proj = Projector(1280, 1120.0 * cols / 640.0)
    return _render(cam, proj, z, seed, device, (70, 64, 59), (steps,) * 3, **common)


def _render(cam, proj, z, seed, device, freqs, steps, ambient, modulation,
            noise, poisson, max_val, dtype, lowmod_frac=0.0, gray=None) -> Pair:
    gen = torch.Generator(device=device).manual_seed(seed + 1000)
    jj = torch.arange(cam.cols, dtype=torch.float64, device=device).view(1, -1)
    up_l = projector_column(cam, proj, jj.expand_as(z), z)
    up_r, _ = right_view(cam, proj, z)
    low_l = low_r = None
    if lowmod_frac > 0:
        low_l = torch.rand(z.shape, generator=gen, device=device) < lowmod_frac
        low_r = torch.rand(z.shape, generator=gen, device=device) < lowmod_frac
    left = fringe_frames(up_l, proj, freqs, steps, ambient, modulation, gen, noise,
                         poisson, max_val, dtype, low_l)
    gen_r = torch.Generator(device=device).manual_seed(seed + 2000)
    right = fringe_frames(up_r, proj, freqs, steps, ambient, modulation, gen_r, noise,
                          poisson, max_val, dtype, low_r)
    if gray is not None:
        period, bits = gray
        left += gray_frames(up_l, proj, period, bits, ambient, ambient + 2 * modulation,
                            gen, noise, max_val, dtype)
        right += gray_frames(up_r, proj, period, bits, ambient, ambient + 2 * modulation,
                             gen_r, noise, max_val, dtype)
    return Pair(cam, proj, left, right, z, tuple(freqs), tuple(steps), q_matrix(cam),
                list(IDENTITY3))


def render_c4(seed: int, device="cpu", rows=1024, cols=1280, f=1600.0, baseline=120.0,
              z0=700.0, noise=True) -> Pair:
    """Differentiation workload: a flat plane target at z0 (slightly tilted,
    seeded) seen by a 1280x1024 pair with the true calibration (f = 1600 px,
    B = 120 mm), three-frequency 4-step fringes, uint16 12-bit."""
    cam = Camera(rows, cols, f, baseline)
    proj = Projector(1280, 1200.0 * cols / 1280.0)
    g = torch.Generator().manual_seed(seed)
    tilt = (torch.rand(2, generator=g, dtype=torch.float64) - 0.5) * 0.2  # mm per px
    ii = torch.arange(rows, dtype=torch.float64, device=device).view(-1, 1)
    jj = torch.arange(cols, dtype=torch.float64, device=device).view(1, -1)
    z = z0 + tilt[0].item() * (jj - cam.cx) + tilt[1].item() * (ii - cam.cy)
    z = z.expand(rows, cols).clone()
    return _render(cam, proj, z, seed, device, (70, 64, 59), (4, 4, 4), ambient=400.0,
                   modulation=750.0, noise=4.0 if noise else 0.0, poisson=noise, max_val=4095,
                   dtype=torch.uint16 if noise else torch.float64)
