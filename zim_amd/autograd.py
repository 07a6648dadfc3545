"""Differentiable triangulation (torch autograd over the HIP adjoint) and the
plane-target calibration refinement it drives.

`triangulate(theta, phL, phR, match_col, nominal)` returns (xyz, valid) with
xyz differentiable with respect to theta (9 calibration parameters: f, cx1,
cy1, cx2, cy2, baseline, wx, wy, wz) and to both raw phase maps. The
autograd node is a torch::autograd::Function in the C++ extension
(csrc/torch_ext.cpp, built by zim_amd._build.build_torch()).

`refine_plane` re-rectifies the raw phase and order maps under the current
calibration on the GPU (psgc_rectify_maps + psgc_remap), re-runs the
order-constrained match (psgc_stereo_match), holds that integer match fixed
and descends the plane-flatness objective along the adjoint gradient with a
backtracking step. A single plane constrains neither scale nor focal length
nor a pure disparity offset (all map planes to planes), so by default only
the principal points and the rectification rotation are free.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import math
import os

import numpy as np
import torch

from . import Context, _capi, default_context

HERE = os.path.dirname(os.path.abspath(__file__))
EXT_DIR = os.path.join(HERE, "_build", "torch")
_ext = None

THETA_NAMES = ("f", "cx1", "cy1", "cx2", "cy2", "baseline", "wx", "wy", "wz")


def _load():
    global _ext
    if _ext is None:
        path = os.path.join(EXT_DIR, "psgc_torch.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built: run zim_amd._build.build_torch()")
        _capi.load()  # libpsgc.so first (the extension links against it)
        spec = importlib.util.spec_from_file_location("psgc_torch", path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _ext = mod
    return _ext


def triangulate(theta, phL, phR, match_col, nominal, ctx: Context | None = None):
    ctx = ctx or default_context()
    stream = _capi.stream_handle(torch.cuda.current_stream(phL.device))
    nominal = torch.as_tensor(nominal, dtype=torch.float64)
    xyz, valid = _load().triangulate(theta, phL, phR, match_col, nominal, ctx.h.value, stream)
    return xyz, valid


def triangulate_views(thetas, phLs, phRs, match_cols, nominal, ctx: Context | None = None):
    """triangulate for many views (pairs) in one forward and one backward
    launch: thetas [V, 9]; per view a raw phase pair and its match columns
    (sizes may differ); nominal [5] shared or [V, 5]. Returns a list of
    (xyz, valid) per view, differentiable like triangulate."""
    ctx = ctx or default_context()
    V = len(phLs)
    stream = _capi.stream_handle(torch.cuda.current_stream(phLs[0].device))
    nom = torch.as_tensor(nominal, dtype=torch.float64)
    if nom.dim() == 1:
        nom = nom.expand(V, 5)
    sizes = torch.tensor([list(p.shape) for p in phLs], dtype=torch.int64)
    L = torch.cat([p.reshape(-1) for p in phLs])
    R = torch.cat([p.reshape(-1) for p in phRs])
    M = torch.cat([m.reshape(-1) for m in match_cols])
    xyz, valid = _load().triangulate_batch(thetas, L, R, M, nom.contiguous(), sizes,
                                           ctx.h.value, stream)
    out, off = [], 0
    for p in phLs:
        n = p.numel()
        out.append((xyz[off:off + n].view(*p.shape, 3), valid[off:off + n].view(*p.shape)))
        off += n
    return out


def point_mask(valid):
    """The points a plane loss sums over: a boolean mask as given (e.g. one
    held fixed across the evaluations of a finite-difference check), or, for
    the triangulation's validity codes, the conditioned points (code 1:
    every phase sample in smoothly unwrapped phase; code 2 marks samples
    across a fringe-order error or an edge, where the point is not a smooth
    function of the calibration)."""
    v = valid.reshape(-1)
    return v if v.dtype == torch.bool else v == 1


def _sel(valid):
    """point_mask as the selection bytes of the plane kernels (1 = counts)."""
    return point_mask(valid).reshape(-1).contiguous().view(torch.uint8)


def _plane_table(xyzs, sels, grads=None):
    tab = (_capi.PlaneView * len(xyzs))()
    for i, (x, m) in enumerate(zip(xyzs, sels)):
        if x.dtype != torch.float64 or not x.is_cuda or x.numel() != 3 * m.numel():
            raise ValueError("plane objectives take f64 device points (n x 3) and one "
                             "selection per point")
        tab[i].xyz = x.data_ptr()
        tab[i].sel = m.data_ptr()
        tab[i].n = m.numel()
        tab[i].grad_xyz = grads[i].data_ptr() if grads is not None else None
    return tab


def _moments(ctx, xyzs, sels):
    """(count [V], mean [V, 3], scatter [V, 3, 3]) as host f64 tensors, from
    psgc_plane_moments_batch (FP64, fixed summation order)."""
    V = len(xyzs)
    cnt, mu, S = (C.c_double * V)(), (C.c_double * (3 * V))(), (C.c_double * (9 * V))()
    tab = _plane_table(xyzs, sels)
    ctx.bind_stream()
    _capi.check(ctx.lib.psgc_plane_moments_batch(ctx.h, tab, V, cnt, mu, S))
    return (torch.tensor(list(cnt), dtype=torch.float64),
            torch.tensor(list(mu), dtype=torch.float64).view(V, 3),
            torch.tensor(list(S), dtype=torch.float64).view(V, 3, 3))


def _plane_grad(ctx, xs, sels, mu, A, b):
    """grad = A (x - mu) + b at the selected points, 0 elsewhere, per view."""
    grads = [torch.empty_like(x) for x in xs]
    # host arrays held by name for the call
    hm, ha, hb = (np.ascontiguousarray(t.numpy(), dtype=np.float64) for t in (mu, A, b))
    ctx.bind_stream()
    _capi.check(ctx.lib.psgc_plane_grad_batch(ctx.h, _plane_table(xs, sels, grads), len(xs),
                                              _capi.ptr(hm), _capi.ptr(ha), _capi.ptr(hb)))
    return grads


class _PlaneScatter(torch.autograd.Function):
    """Scatter matrices [V, 3, 3] (host) of the selected points of each view
    about their mean, differentiable in the points: for dL/dS = G the
    gradient is (G + G^T)(x - mean) at the selected points (the mean's own
    terms sum to zero over them)."""

    @staticmethod
    def forward(actx, ctx, sels, *xyzs):
        xs = [x.detach().reshape(-1, 3).contiguous() for x in xyzs]
        _, mu, S = _moments(ctx, xs, sels)
        actx.ctx, actx.sels, actx.mu = ctx, sels, mu
        actx.shapes = [x.shape for x in xyzs]
        actx.save_for_backward(*xs)
        return S

    @staticmethod
    def backward(actx, gS):
        xs = actx.saved_tensors
        G = gS.detach().to("cpu", torch.float64)
        A = G + G.transpose(1, 2)
        b = torch.zeros(len(xs), 3, dtype=torch.float64)
        grads = _plane_grad(actx.ctx, xs, actx.sels, actx.mu, A, b)
        return (None, None, *[g.view(sh) for g, sh in zip(grads, actx.shapes)])


class _PlaneLoss(torch.autograd.Function):
    """Per view the sum of squared residuals r = n . (x - mean) of the
    selected points to their least-squares plane (n: the scatter matrix's
    eigenvector of the smallest eigenvalue, held fixed: the loss is
    stationary in it). Gradient 2 n (r - sum(r) / N) at the selected points,
    the second term from the mean."""

    @staticmethod
    def forward(actx, ctx, sels, *xyzs):
        xs = [x.detach().reshape(-1, 3).contiguous() for x in xyzs]
        V = len(xs)
        cnt, mu, S = _moments(ctx, xs, sels)
        n = torch.linalg.eigh(S)[1][:, :, 0].contiguous()
        sq, sm = (C.c_double * V)(), (C.c_double * V)()
        hm, hn = (np.ascontiguousarray(t.numpy(), dtype=np.float64) for t in (mu, n))
        ctx.bind_stream()
        _capi.check(ctx.lib.psgc_plane_residuals_batch(ctx.h, _plane_table(xs, sels), V,
                                                       _capi.ptr(hm), _capi.ptr(hn), sq, sm))
        actx.ctx, actx.sels, actx.mu, actx.n = ctx, sels, mu, n
        actx.rbar = torch.tensor(list(sm), dtype=torch.float64) / cnt.clamp_min(1.0)
        actx.shapes = [x.shape for x in xyzs]
        actx.save_for_backward(*xs)
        return torch.tensor(list(sq), dtype=torch.float64)

    @staticmethod
    def backward(actx, g):
        xs = actx.saved_tensors
        g = g.detach().to("cpu", torch.float64)
        n = actx.n
        A = 2.0 * g[:, None, None] * n[:, :, None] * n[:, None, :]
        b = -2.0 * (g * actx.rbar)[:, None] * n
        grads = _plane_grad(actx.ctx, xs, actx.sels, actx.mu, A, b)
        return (None, None, *[gr.view(sh) for gr, sh in zip(grads, actx.shapes)])


def _scatter_views(views, ctx=None):
    ctx = ctx or default_context()
    return _PlaneScatter.apply(ctx, [_sel(v) for _, v in views], *[x for x, _ in views])


def plane_losses(views, ctx: Context | None = None):
    """Sum of squared residuals of the points of each (xyz, valid) (see
    point_mask) to their least-squares plane, summed from the residuals
    n . (x - mean) themselves, n the plane normal: the smallest eigenvalue
    read off the scatter matrix carries an absolute error of ~1e-16 times its
    largest eigenvalue (the plane's extent squared, times the point count),
    which swamps finite differences of the loss. The loss is stationary in n,
    so n needs no gradient. Moments, residuals and gradient run in the
    library's fused kernels (psgc_plane_*_batch: one launch per pass for all
    views, FP64 sums in a fixed order); the 3 x 3 eigenproblems of all views
    are solved on the host. Returns a host tensor [V]."""
    ctx = ctx or default_context()
    return _PlaneLoss.apply(ctx, [_sel(v) for _, v in views], *[x for x, _ in views])


def plane_loss(xyz, valid, ctx: Context | None = None):
    """plane_losses of one view."""
    return plane_losses([(xyz, valid)], ctx)[0]


def plane_flatnesses(views, ctx: Context | None = None):
    """Scale-invariant flatness of each (xyz, valid): smallest eigenvalue over
    the trace of the scatter matrix of the valid points (host eigensolve, one
    round trip for all views). Returns a host tensor."""
    ev = torch.linalg.eigvalsh(_scatter_views(views, ctx))
    return ev[:, 0] / ev.sum(1)


def plane_flatness(xyz, valid, ctx: Context | None = None):
    """plane_flatnesses of one view."""
    return plane_flatnesses([(xyz, valid)], ctx)[0]


def default_scale(theta0):
    """Perturbation scale of each parameter: 1 % focal, 3 px principal
    points, 0.5 % baseline, 0.05 deg rotations."""
    t = torch.as_tensor(theta0, dtype=torch.float64)
    return torch.tensor([0.01 * float(t[0]), 3.0, 3.0, 3.0, 3.0, 0.005 * float(t[5]),
                         math.radians(0.05), math.radians(0.05), math.radians(0.05)],
                        dtype=torch.float64)


def _rot(wx, wy, wz):
    cx, sx, cy, sy, cz, sz = (math.cos(wx), math.sin(wx), math.cos(wy), math.sin(wy),
                              math.cos(wz), math.sin(wz))
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def rectified_match(theta, nominal, phL, phR, kL, kR, pha_f, grad_threshold=1.0,
                    ctx: Context | None = None):
    """Rectifies the raw maps under theta on the GPU (bilinear phase,
    nearest order) and returns the integer match column per left pixel."""
    ctx = (ctx or default_context()).bind_stream()
    lib = ctx.lib
    f, cx1, cy1, cx2, cy2, B, wx, wy, wz = [float(v) for v in theta]
    f0, cLx, cLy, cRx, cRy = [float(v) for v in nominal]
    rows, cols = phL.shape
    dev = phL.device
    K = lambda ff, a, b: np.array([[ff, 0, a], [0, ff, b], [0, 0, 1.0]])  # noqa: E731
    out = []
    for Kraw, R, Krect, ph, k in ((K(f0, cLx, cLy), np.eye(3), K(f, cx1, cy1), phL, kL),
                                  (K(f0, cRx, cRy), _rot(wx, wy, wz).T, K(f, cx2, cy2), phR, kR)):
        mx = torch.empty((rows, cols), dtype=torch.float32, device=dev)
        my = torch.empty_like(mx)
        _capi.check(lib.psgc_rectify_maps(ctx.h, _capi.ptr(np.ascontiguousarray(Kraw)),
                                          _capi.ptr(np.ascontiguousarray(R)),
                                          _capi.ptr(np.ascontiguousarray(Krect)), rows, cols,
                                          _capi.ptr(mx), _capi.ptr(my), _capi.MEM_DEVICE))
        phr = torch.empty_like(ph)
        kr = torch.empty_like(k)
        _capi.check(lib.psgc_remap(ctx.h, 0, _capi.ptr(ph), rows, cols, _capi.ptr(mx),
                                   _capi.ptr(my), rows, cols, _capi.ptr(phr), _capi.MEM_DEVICE))
        _capi.check(lib.psgc_remap(ctx.h, 1, _capi.ptr(k), rows, cols, _capi.ptr(mx),
                                   _capi.ptr(my), rows, cols, _capi.ptr(kr), _capi.MEM_DEVICE))
        out.append((phr, kr))
    disp = torch.empty((rows, cols), dtype=torch.float64, device=dev)
    mc = torch.empty((rows, cols), dtype=torch.int32, device=dev)
    (pl, kl), (pr, kr) = out
    _capi.check(lib.psgc_stereo_match(ctx.h, _capi.ptr(pl), _capi.ptr(pr), _capi.ptr(kl),
                                      _capi.ptr(kr), rows, cols, int(pha_f), grad_threshold,
                                      _capi.ptr(disp), _capi.ptr(mc), _capi.MEM_DEVICE))
    return mc


def refine_plane(theta_start, theta_ref, phL, phR, kL, kR, nominal, pha_f, outer=6, inner=6,
                 step=0.5, free=(1, 2, 3, 4, 6, 7, 8), ctx: Context | None = None):
    """Plane-target refinement by backtracking gradient descent in the
    normalized coordinates z = (theta - theta_ref) / default_scale. Returns
    (theta, history of (flatness before, flatness after) per outer step)."""
    ctx = ctx or default_context()
    scale = default_scale(theta_ref)
    ref = torch.as_tensor(theta_ref, dtype=torch.float64)
    mask = torch.zeros(9, dtype=torch.float64)
    mask[list(free)] = 1.0
    z = ((torch.as_tensor(theta_start, dtype=torch.float64) - ref) / scale).clone()
    hist = []

    def obj(zz, mc, grad=False):
        th = (ref + zz * scale).requires_grad_(grad)
        v = plane_flatness(*triangulate(th, phL, phR, mc, nominal, ctx), ctx)
        if grad:
            v.backward()
            return v.detach(), th.grad * scale * mask
        return v.detach()

    for _ in range(outer):
        mc = rectified_match(ref + z * scale, nominal, phL, phR, kL, kR, pha_f, ctx=ctx)
        start = None
        for _ in range(inner):
            f0, g = obj(z, mc, grad=True)
            if start is None:
                start = float(f0)
            gn = float(g.norm())
            if gn == 0.0:
                break
            d = -g / gn
            a = step
            while a > 1e-4:
                f1 = obj(z + a * d, mc)
                if f1 < f0:
                    z = z + a * d
                    break
                a *= 0.5
            else:
                break
        hist.append((start, float(obj(z, mc))))
    return ref + z * scale, hist


def refine_planes(thetas_start, thetas_ref, phLs, phRs, kLs, kRs, nominal, pha_f, outer=6,
                  inner=6, step=0.5, free=(1, 2, 3, 4, 6, 7, 8), ctx: Context | None = None):
    """refine_plane for many views at once (nominal [5] shared or [V, 5]):
    each inner step is one batched
    forward + backward over all views (their flatness objectives are
    separate, so one summed backward yields every view's gradient), and the
    backtracking line search runs for all views together (a trial is one
    batched forward). thetas [V, 9]. Returns (thetas, per view the history
    of (flatness before, flatness after) per outer step)."""
    ctx = ctx or default_context()
    V = len(phLs)
    nom = torch.as_tensor(nominal, dtype=torch.float64)
    if nom.dim() == 1:
        nom = nom.expand(V, 5)
    ref = torch.as_tensor(thetas_ref, dtype=torch.float64).reshape(V, 9)
    scale = torch.stack([default_scale(r) for r in ref])
    mask = torch.zeros(9, dtype=torch.float64)
    mask[list(free)] = 1.0
    z = ((torch.as_tensor(thetas_start, dtype=torch.float64).reshape(V, 9) - ref) / scale).clone()
    hist = [[] for _ in range(V)]

    def objs(zz, mcs, grad=False):
        th = (ref + zz * scale).requires_grad_(grad)
        res = triangulate_views(th, phLs, phRs, mcs, nom, ctx)
        vals = plane_flatnesses(res, ctx)
        if grad:
            vals.sum().backward()
            return vals.detach().cpu(), th.grad * scale * mask
        return vals.detach().cpu()

    for _ in range(outer):
        th_now = ref + z * scale
        mcs = [rectified_match(th_now[v], nom[v], phLs[v], phRs[v], kLs[v], kRs[v], pha_f,
                               ctx=ctx) for v in range(V)]
        start = None
        for _ in range(inner):
            f0, g = objs(z, mcs, grad=True)
            if start is None:
                start = f0.clone()
            gn = g.norm(dim=1, keepdim=True)
            d = torch.where(gn > 0, -g / gn.clamp_min(1e-300), torch.zeros_like(g))
            a = torch.full((V, 1), step, dtype=torch.float64)
            pending = gn[:, 0] > 0
            while bool(pending.any()):  # every view's step halves to 1e-4 at most
                trial = torch.where(pending[:, None], z + a * d, z)
                f1 = objs(trial, mcs)
                better = pending & (f1 < f0)
                z = torch.where(better[:, None], trial, z)
                pending = pending & ~better
                a = torch.where(pending[:, None], a * 0.5, a)
                pending = pending & (a[:, 0] > 1e-4)
            if not bool((gn[:, 0] > 0).any()):
                break
        end = objs(z, mcs)
        for v in range(V):
            hist[v].append((float(start[v]), float(end[v])))
    return ref + z * scale, hist
