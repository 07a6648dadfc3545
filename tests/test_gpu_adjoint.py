"""Differentiable triangulation: the HIP forward against an independent FP64
torch implementation (tests/adjoint_ref.py), the hand-written adjoint against
torch autograd of that implementation (FP64 reference gradient) and against
central finite differences (h = 1e-4), through the libtorch
autograd::Function, at 256 x 320 and at the full 1280 x 1024; and the
plane-target calibration refinement it drives.

Scene: a flat plane target at 700 mm rendered with the true calibration
(f = 1600 px, B = 120 mm), 3-frequency 4-step uint16 fringes; the pair is
reconstructed with a seeded perturbation of focal (+-1 %), baseline
(+-0.5 %), principal points (+-3 px) and rectification rotation (+-0.05 deg).
"""
import functools
import math

import numpy as np
import pytest
import torch

import zim_amd
from zim_amd import autograd as zag
from zim_amd import synth
from tests import adjoint_ref, scenes

pytestmark = pytest.mark.gpu


_ORDERS = {}   # fringe orders of each rendered plane pair, by (rows, cols, seed)


@functools.lru_cache(maxsize=None)
def plane_pair(rows=256, cols=320, seed=4):
    f = 1600.0 * cols / 1280.0
    p = synth.render_c4(seed, device="cuda", rows=rows, cols=cols, f=f)
    ctx = zim_amd.Context(0)
    dcfg = scenes.c2_decode_config()
    L = zim_amd.decode(dcfg, p.left, ctx)
    R = zim_amd.decode(dcfg, p.right, ctx)
    rcfg = zim_amd.ReconConfig.default_gpu(use_mls=0, outlier_mode=0)
    _, _, _, mc = zim_amd.reconstruct(rcfg, L[0], R[0], L[1], R[1], 70, p.Q, p.R1, ctx)
    torch.cuda.synchronize()
    cam = p.cam
    theta0 = torch.tensor([cam.f, cam.cx, cam.cy, cam.cx, cam.cy, cam.baseline, 0, 0, 0],
                          dtype=torch.float64)
    nominal = torch.tensor([cam.f, cam.cx, cam.cy, cam.cx, cam.cy], dtype=torch.float64)
    _ORDERS[(rows, cols, seed)] = (L[1].contiguous(), R[1].contiguous())
    return ctx, L[0].contiguous(), R[0].contiguous(), mc.contiguous(), theta0, nominal


def plane_orders(rows=256, cols=320, seed=4):
    """Fringe orders (left, right) of plane_pair(rows, cols, seed); cached
    with it, so they belong to the same pair whatever was called between."""
    plane_pair(rows, cols, seed)
    return _ORDERS[(rows, cols, seed)]


def perturbed(theta0, seed=7):
    g = np.random.default_rng(seed)
    d = np.zeros(9)
    d[0] = theta0[0].item() * 0.01 * g.uniform(-1, 1)
    d[1:5] = 3.0 * g.uniform(-1, 1, 4)
    d[5] = theta0[5].item() * 0.005 * g.uniform(-1, 1)
    d[6:9] = math.radians(0.05) * g.uniform(-1, 1, 3)
    return theta0 + torch.from_numpy(d)


def _codes(ok, cond):
    """The triangulation's validity codes from the reference's masks."""
    return np.where(ok.numpy(), np.where(cond.numpy(), 1, 2), 0).astype(np.uint8)


def test_forward_matches_fp64_reference():
    ctx, phL, phR, mc, theta0, nominal = plane_pair()
    for th in (theta0, perturbed(theta0)):
        xyz, valid = zag.triangulate(th, phL, phR, mc, nominal, ctx)
        rx, rok, rcond = adjoint_ref.forward(th, nominal, phL.cpu(), phR.cpu(), mc.cpu())
        assert (valid == 1).sum().item() > 0.3 * valid.numel()
        assert np.array_equal(valid.cpu().numpy(), _codes(rok, rcond))
        assert (xyz.cpu() - rx).abs().max().item() < 1e-8


def test_forward_at_true_calibration_equals_pipeline_disparity():
    """At theta0 the remaps are the identity, so the triangulation equals
    the Q reprojection of the pipeline's own f64 disparity for the same
    matches: Z = f B / d."""
    import ctypes as C
    from zim_amd import _capi
    ctx, phL, phR, mc, theta0, nominal = plane_pair()
    kL, kR = plane_orders()
    xyz, valid = zag.triangulate(theta0, phL, phR, mc, nominal, ctx)
    rows, cols = phL.shape
    disp = torch.empty((rows, cols), dtype=torch.float64, device="cuda")
    mc2 = torch.empty((rows, cols), dtype=torch.int32, device="cuda")
    ctx.bind_stream()
    _capi.check(ctx.lib.psgc_stereo_match(ctx.h, _capi.ptr(phL), _capi.ptr(phR), _capi.ptr(kL),
                                          _capi.ptr(kR), rows, cols, 70, 1.0, _capi.ptr(disp),
                                          _capi.ptr(mc2), _capi.MEM_DEVICE))
    torch.cuda.synchronize()
    assert torch.equal(mc2, mc)
    m = valid.bool() & (mc >= 0)
    Z = theta0[0].item() * theta0[5].item() / disp[m]
    assert (xyz[..., 2][m] - Z).abs().max().item() < 1e-6 * Z.abs().max().item()


def test_adjoint_matches_autograd_reference():
    ctx, phL, phR, mc, theta0, nominal = plane_pair()
    th = perturbed(theta0)
    g = torch.Generator().manual_seed(0)
    w = torch.randn(phL.shape + (3,), generator=g, dtype=torch.float64)
    # GPU adjoint through the autograd Function
    t1 = th.clone().requires_grad_(True)
    pl = phL.clone().requires_grad_(True)
    pr = phR.clone().requires_grad_(True)
    xyz, valid = zag.triangulate(t1, pl, pr, mc, nominal, ctx)
    (xyz * w.cuda()).sum().backward()
    # FP64 reference gradient
    t2 = th.clone().requires_grad_(True)
    pl2 = phL.cpu().clone().requires_grad_(True)
    pr2 = phR.cpu().clone().requires_grad_(True)
    rx, rok, _ = adjoint_ref.forward(t2, nominal, pl2, pr2, mc.cpu())
    (rx * w).sum().backward()
    rel = ((t1.grad - t2.grad).abs() / t2.grad.abs().clamp_min(1e-12)).max().item()
    assert rel < 1e-6, (t1.grad, t2.grad)
    for a, b in ((pl.grad.cpu(), pl2.grad), (pr.grad.cpu(), pr2.grad)):
        assert (a - b).abs().max().item() <= 1e-9 * max(1.0, b.abs().max().item())


def _fd_check(rows, cols, seed, h=1e-4):
    """Adjoint of the plane loss against central differences with step h in
    the normalized calibration coordinates z = (theta - theta0) / scale
    (scale: 1 % focal, 3 px principal points, 0.5 % baseline, 0.05 deg
    rotations). The loss sums over a fixed set of points, as the integer
    match is fixed: those conditioned at theta and triangulated at every
    evaluation point. Returns the relative error per parameter, against
    max(|gradient|, 1e-3 of the largest)."""
    ctx, phL, phR, mc, theta0, nominal = plane_pair(rows, cols)
    scale = zag.default_scale(theta0)
    thp = perturbed(theta0, seed)
    steps = []
    for k in range(9):
        for sgn in (1.0, -1.0):
            e = torch.zeros(9, dtype=torch.float64)
            e[k] = sgn * h * scale[k]
            steps.append(e)
    with torch.no_grad():
        M = zag.point_mask(zag.triangulate(thp, phL, phR, mc, nominal, ctx)[1])
        for e in steps:
            M = M & (zag.triangulate(thp + e, phL, phR, mc, nominal, ctx)[1].reshape(-1) != 0)
    th = thp.clone().requires_grad_(True)
    zag.plane_loss(zag.triangulate(th, phL, phR, mc, nominal, ctx)[0], M).backward()
    grad = th.grad.clone() * scale
    fd = torch.zeros(9, dtype=torch.float64)
    with torch.no_grad():
        for k in range(9):
            lp = zag.plane_loss(zag.triangulate(thp + steps[2 * k], phL, phR, mc, nominal, ctx)[0], M)
            lm = zag.plane_loss(zag.triangulate(thp + steps[2 * k + 1], phL, phR, mc, nominal,
                                                ctx)[0], M)
            fd[k] = (lp - lm) / (2 * h)
    rel = (grad - fd).abs() / grad.abs().clamp_min(1e-3 * grad.abs().max())
    print(f"{rows}x{cols} seed {seed}: points {int(M.sum())}, rel", rel.tolist(), flush=True)
    return rel


def test_plane_loss_adjoint_vs_central_differences():
    assert _fd_check(256, 320, 7).max().item() <= 1e-4


@pytest.mark.parametrize("seed", [7, 8])
def test_full_size_adjoint_vs_central_differences(seed):
    """At the differentiation workload's full 1280 x 1024 size, h = 1e-4:
    every parameter within 1e-4. The triangulation resamples with C1
    Catmull-Rom cubics (no kinks at pixel boundaries for the stencil to
    straddle), and the loss leaves out points whose samples span a phase
    jump of pi or more (at this scene's right edge the projector's last
    fringe decodes to wrong orders; a sample there makes its point tens of
    millimetres off the plane and its derivative meaningless)."""
    assert _fd_check(1024, 1280, seed).max().item() <= 1e-4


def test_full_size_adjoint_matches_fp64_reference_gradient():
    """The adjoint's theta gradient at 1280 x 1024 against torch autograd of
    the FP64 reference model (tests/adjoint_ref.py, on the host), for a
    seeded random weighting of every coordinate."""
    ctx, phL, phR, mc, theta0, nominal = plane_pair(1024, 1280)
    th = perturbed(theta0, 7)
    g = torch.Generator().manual_seed(3)
    w = torch.randn(phL.shape + (3,), generator=g, dtype=torch.float64)
    t1 = th.clone().requires_grad_(True)
    xyz, valid = zag.triangulate(t1, phL, phR, mc, nominal, ctx)
    (xyz * w.cuda()).sum().backward()
    print("[adjoint] FP64 reference gradient at 1280x1024 on the host", flush=True)
    t2 = th.clone().requires_grad_(True)
    rx, rok, rcond = adjoint_ref.forward(t2, nominal, phL.cpu(), phR.cpu(), mc.cpu())
    (rx * w).sum().backward()
    assert np.array_equal(valid.cpu().numpy(), _codes(rok, rcond))
    rel = ((t1.grad - t2.grad).abs() / t2.grad.abs().clamp_min(1e-12)).max().item()
    print("max relative difference", rel)
    assert rel < 1e-6, (t1.grad, t2.grad)


def test_plane_refinement_reduces_flatness():
    """Refinement from the perturbed calibration: each outer step re-rectifies
    and re-matches on the GPU, then descends the flatness objective along
    the adjoint gradient; the objective must decrease."""
    ctx, phL, phR, mc, theta0, nominal = plane_pair()
    kL, kR = plane_orders()
    th = perturbed(theta0)
    theta, hist = zag.refine_plane(th, theta0, phL, phR, kL, kR, nominal, 70, outer=4, inner=4,
                                   ctx=ctx)
    print("flatness per outer step (start, end):", hist)
    assert all(b <= a for a, b in hist)
    assert hist[-1][1] < hist[0][0]
    assert torch.isfinite(theta).all()


# ------------------------------------------------------- many views at once --

def _views():
    """Three plane-target pairs of different sizes and seeds with their
    decoded phases and matches at the true calibration."""
    out = []
    for rows, cols, seed in ((256, 320, 4), (200, 288, 5), (240, 256, 6)):
        f = 1600.0 * cols / 1280.0
        p = synth.render_c4(seed, device="cuda", rows=rows, cols=cols, f=f)
        ctx = zim_amd.Context(0)
        dcfg = scenes.c2_decode_config()
        L = zim_amd.decode(dcfg, p.left, ctx)
        R = zim_amd.decode(dcfg, p.right, ctx)
        rcfg = zim_amd.ReconConfig.default_gpu(use_mls=0, outlier_mode=0)
        _, _, _, mc = zim_amd.reconstruct(rcfg, L[0], R[0], L[1], R[1], 70, p.Q, p.R1, ctx)
        torch.cuda.synchronize()
        cam = p.cam
        theta0 = torch.tensor([cam.f, cam.cx, cam.cy, cam.cx, cam.cy, cam.baseline, 0, 0, 0],
                              dtype=torch.float64)
        nominal = torch.tensor([cam.f, cam.cx, cam.cy, cam.cx, cam.cy], dtype=torch.float64)
        out.append(dict(phL=L[0].contiguous(), phR=R[0].contiguous(), kL=L[1].contiguous(),
                        kR=R[1].contiguous(), mc=mc.contiguous(), theta0=theta0,
                        nominal=nominal))
    return out


def test_batched_views_equal_per_view_calls():
    """One forward and one backward launch over three views of different
    sizes give each view's per-view results: xyz bit-identical, gradients
    equal up to the order of the f64 atomic sums."""
    ctx = zim_amd.Context(0)
    vs = _views()
    thetas = torch.stack([perturbed(v["theta0"], seed=11 + i) for i, v in enumerate(vs)])
    noms = torch.stack([v["nominal"] for v in vs])
    g = torch.Generator().manual_seed(1)
    ws = [torch.randn(v["phL"].shape + (3,), generator=g, dtype=torch.float64).cuda() for v in vs]
    tb = thetas.clone().requires_grad_(True)
    pls = [v["phL"].clone().requires_grad_(True) for v in vs]
    res = zag.triangulate_views(tb, pls, [v["phR"] for v in vs], [v["mc"] for v in vs], noms, ctx)
    sum((x * w).sum() for (x, _), w in zip(res, ws)).backward()
    for i, v in enumerate(vs):
        t1 = thetas[i].clone().requires_grad_(True)
        pl = v["phL"].clone().requires_grad_(True)
        x1, ok1 = zag.triangulate(t1, pl, v["phR"], v["mc"], v["nominal"], ctx)
        (x1 * ws[i]).sum().backward()
        assert torch.equal(res[i][0].detach(), x1.detach())
        assert torch.equal(res[i][1], ok1)
        assert ok1.float().mean() > 0.3
        rel = ((tb.grad[i] - t1.grad).abs() / t1.grad.abs().clamp_min(1e-12)).max().item()
        assert rel < 1e-9, (tb.grad[i], t1.grad)
        assert (pls[i].grad - pl.grad).abs().max().item() <= 1e-12 * max(1.0, pl.grad.abs().max().item())


def test_batched_plane_refinement_reduces_flatness():
    """Three views refined together (one batched adjoint launch per inner
    step): every view's flatness decreases."""
    ctx = zim_amd.Context(0)
    vs = _views()
    thetas = torch.stack([perturbed(v["theta0"], seed=21 + i) for i, v in enumerate(vs)])
    refs = torch.stack([v["theta0"] for v in vs])
    th, hist = zag.refine_planes(thetas, refs, [v["phL"] for v in vs], [v["phR"] for v in vs],
                                 [v["kL"] for v in vs], [v["kR"] for v in vs],
                                 torch.stack([v["nominal"] for v in vs]), 70, outer=3, inner=3,
                                 ctx=ctx)
    print("per-view flatness (start, end) per outer step:", hist)
    for h in hist:
        assert all(b <= a for a, b in h)
        assert h[-1][1] < h[0][0]
    assert torch.isfinite(th).all()
