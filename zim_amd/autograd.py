"""Differentiable triangulation (torch autograd over the HIP adjoint) and the
plane-target calibration refinement it drives.

`triangulate(theta, phL, phR, match_col, nominal)` returns (xyz, valid) with
xyz differentiable with respect to theta (9 calibration parameters: f, cx1,
cy1, cx2, cy2, baseline, wx, wy, wz) and to both raw phase maps. The
autograd node is a torch::autograd::Function in the C++ extension
(csrc/torch_ext.cpp, built by zim_amd._build.build_torch()).
"""
from __future__ import annotations

import importlib.util
import os

import torch

from . import Context, default_context

HERE = os.path.dirname(os.path.abspath(__file__))
EXT_DIR = os.path.join(HERE, "_build", "torch")
_ext = None


def _load():
    global _ext
    if _ext is None:
        path = os.path.join(EXT_DIR, "psgc_torch.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built: run zim_amd._build.build_torch()")
        from . import _capi
        _capi.load()  # libpsgc.so first (the extension links against it)
        spec = importlib.util.spec_from_file_location("psgc_torch", path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _ext = mod
    return _ext


def triangulate(theta, phL, phR, match_col, nominal, ctx: Context | None = None):
    ctx = ctx or default_context()
    stream = torch.cuda.current_stream(phL.device).cuda_stream
    nominal = torch.as_tensor(nominal, dtype=torch.float64)
    xyz, valid = _load().triangulate(theta, phL, phR, match_col, nominal, ctx.h.value, stream)
    return xyz, valid


def plane_loss(xyz, valid):
    """Sum of squared residuals of the valid points to their least-squares
    plane: the smallest eigenvalue of the scatter matrix."""
    m = valid.bool()
    pts = xyz[m]
    e = pts - pts.mean(0)
    return torch.linalg.eigvalsh(e.T @ e)[0]


THETA_NAMES = ("f", "cx1", "cy1", "cx2", "cy2", "baseline", "wx", "wy", "wz")


def refine_plane(theta0, phL, phR, match_col, nominal, iters=60, lr=None, ctx=None,
                 free=None):
    """Plane-target calibration refinement: minimises the plane-fit residual
    of the reconstructed target over the free calibration parameters by
    gradient descent with per-parameter scales (Adam). Returns (theta,
    history of losses)."""
    scale = torch.tensor([10.0, 1.0, 1.0, 1.0, 1.0, 0.5, 1e-3, 1e-3, 1e-3], dtype=torch.float64)
    lr = lr if lr is not None else 0.05
    free = free if free is not None else list(range(9))
    mask = torch.zeros(9, dtype=torch.float64)
    mask[free] = 1.0
    theta0 = torch.as_tensor(theta0, dtype=torch.float64)
    z = torch.zeros(9, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([z], lr=lr)
    hist = []
    for _ in range(iters):
        opt.zero_grad()
        theta = theta0 + z * scale * mask
        xyz, valid = triangulate(theta, phL, phR, match_col, nominal, ctx)
        loss = plane_loss(xyz, valid)
        loss.backward()
        opt.step()
        hist.append(float(loss.detach()))
    return (theta0 + z.detach() * scale * mask), hist
