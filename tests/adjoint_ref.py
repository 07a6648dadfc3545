"""FP64 reference of the differentiable rectification + sub-pixel root + Q
triangulation, written with torch tensor ops so that torch's own autograd
gives the reference gradient for the hand-written HIP adjoint.

theta = (f, cx1, cy1, cx2, cy2, B, wx, wy, wz); nominal = (f0, c0Lx, c0Ly,
c0Rx, c0Ry). Catmull-Rom bicubic sampling over 4 x 4 taps (the sample
needs u in [1, cols - 2], v in [1, rows - 2]); a point is conditioned when
every sample's adjacent taps differ by less than pi (Itoh's condition)."""
from __future__ import annotations

import torch


def _rot(wx, wy, wz):
    one = torch.ones_like(wx)
    zero = torch.zeros_like(wx)
    cx, sx, cy, sy, cz, sz = wx.cos(), wx.sin(), wy.cos(), wy.sin(), wz.cos(), wz.sin()
    Rx = torch.stack([torch.stack([one, zero, zero]), torch.stack([zero, cx, -sx]),
                      torch.stack([zero, sx, cx])])
    Ry = torch.stack([torch.stack([cy, zero, sy]), torch.stack([zero, one, zero]),
                      torch.stack([-sy, zero, cy])])
    Rz = torch.stack([torch.stack([cz, -sz, zero]), torch.stack([sz, cz, zero]),
                      torch.stack([zero, zero, one])])
    return Rz @ Ry @ Rx


def _catmull_rom(t):
    t2, t3 = t * t, t * t * t
    return ((-t3 + 2 * t2 - t) / 2, (3 * t3 - 5 * t2 + 2) / 2, (-3 * t3 + 4 * t2 + t) / 2,
            (t3 - t2) / 2)


def _bicubic(m, u, v):
    """Catmull-Rom bicubic sample of m at (u, v) over the 4 x 4 taps around
    it: (value, taps inside the map, adjacent taps pairwise within pi)."""
    rows, cols = m.shape
    ok = (u >= 1) & (u <= cols - 2) & (v >= 1) & (v <= rows - 2) & (rows >= 4) & (cols >= 4)
    x0 = torch.clamp(torch.floor(u.detach()), 1, max(cols - 3, 1)).long()
    y0 = torch.clamp(torch.floor(v.detach()), 1, max(rows - 3, 1)).long()
    wx, wy = _catmull_rom(u - x0.double()), _catmull_rom(v - y0.double())
    flat = m.reshape(-1)
    g = lambda yy, xx: flat[(yy * cols + xx).clamp(0, rows * cols - 1)]  # noqa: E731
    taps = [[g(y0 + j - 1, x0 + i - 1) for i in range(4)] for j in range(4)]
    val = 0
    for j in range(4):
        for i in range(4):
            val = val + wy[j] * wx[i] * taps[j][i]
    pi = 3.14159265358979323846
    smooth = torch.ones_like(ok)
    for j in range(4):
        for i in range(4):
            if i > 0:
                smooth = smooth & ((taps[j][i] - taps[j][i - 1]).abs() < pi)
            if j > 0:
                smooth = smooth & ((taps[j][i] - taps[j - 1][i]).abs() < pi)
    return val, ok, smooth.detach()


def forward(theta, nominal, phL, phR, match_col):
    """Returns xyz (rows, cols, 3), the validity mask and the conditioning
    mask (every sample's adjacent taps within pi)."""
    f, cx1, cy1, cx2, cy2, B, wx, wy, wz = [theta[k] for k in range(9)]
    f0, cLx, cLy, cRx, cRy = [float(v) for v in nominal]
    rows, cols = phL.shape
    ii, jj = torch.meshgrid(torch.arange(rows, dtype=torch.float64),
                            torch.arange(cols, dtype=torch.float64), indexing="ij")
    b = torch.as_tensor(match_col).long()
    ok = (b >= 1) & (b <= cols - 2)
    ul = f0 / f * (jj - cx1) + cLx
    vl = f0 / f * (ii - cy1) + cLy
    pl, okl, sm = _bicubic(phL, ul, vl)
    ok = ok & okl
    R = _rot(wx, wy, wz)
    cs = []
    for m in range(3):
        x = (b - 1 + m).double()
        r0 = (x - cx2) / f
        r1 = (ii - cy2) / f
        q0 = R[0, 0] * r0 + R[0, 1] * r1 + R[0, 2]
        q1 = R[1, 0] * r0 + R[1, 1] * r1 + R[1, 2]
        q2 = R[2, 0] * r0 + R[2, 1] * r1 + R[2, 2]
        u = f0 * q0 / q2 + cRx
        v = f0 * q1 / q2 + cRy
        pr, okr, smr = _bicubic(phR, u, v)
        ok = ok & okr
        sm = sm & smr
        cs.append(pr - pl)
    c0, c1, c2 = cs
    N = c1 * (c0 - c2)
    Dn = 2.0 * (c0 - c1) * (c1 - c2)
    ok = ok & (Dn.abs() > 1e-9)
    delta = torch.where(ok, N / torch.where(ok, Dn, torch.ones_like(Dn)), torch.zeros_like(N))
    d = jj - b.double() - delta
    D = d - (cx1 - cx2)
    ok = ok & (D > 0)
    Ds = torch.where(ok, D, torch.ones_like(D))
    X = (jj - cx1) * B / Ds
    Y = (ii - cy1) * B / Ds
    Z = f * B / Ds
    xyz = torch.stack([X, Y, Z], -1) * ok[..., None]
    return xyz, ok, ok & sm


def plane_loss(xyz, ok):
    """Sum of squared residuals to the least-squares plane (smallest
    eigenvalue of the scatter matrix of the valid points)."""
    pts = xyz[ok]
    c = pts.mean(0)
    e = pts - c
    S = e.T @ e
    return torch.linalg.eigvalsh(S)[0]
