"""The fused plane-fit objectives of the calibration refinement
(psgc_plane_moments/residuals/grad_batch behind zim_amd.autograd's
plane_losses and plane_flatnesses) against FP64 references of the same
mathematics on the host: numpy moments, and torch autograd of the masked
residual and eigenvalue formulations."""
import numpy as np
import pytest
import torch

from zim_amd import autograd as zag

gpu = pytest.mark.gpu


def _views(seed=0):
    """Three views of different sizes: noisy tilted planes with outliers,
    selections as codes (0/1/2) or booleans."""
    g = np.random.default_rng(seed)
    out = []
    for i, n in enumerate((1000, 77777, 4096 * 3 + 5)):
        uv = g.uniform(-60, 60, (n, 2))
        z = 700 + 0.3 * uv[:, 0] - 0.1 * uv[:, 1] + g.normal(0, 0.05, n)
        x = np.column_stack([uv, z])
        x[g.random(n) < 0.01] += g.normal(0, 30, (1, 3))
        if i == 1:
            sel = torch.from_numpy(g.integers(0, 3, n).astype(np.uint8))
        else:
            sel = torch.from_numpy(g.random(n) < 0.9)
        out.append((torch.from_numpy(x).cuda(), sel.cuda()))
    return out


def _ref_moments(x, m):
    p = x[m]
    mu = p.mean(0)
    d = p - mu
    return len(p), mu, d.T @ d


@gpu
def test_moments_match_fp64_numpy(ctx):
    views = _views()
    xs = [x.reshape(-1, 3).contiguous() for x, _ in views]
    sels = [zag._sel(v) for _, v in views]
    cnt, mu, S = zag._moments(ctx, xs, sels)
    for k, (x, v) in enumerate(views):
        m = zag.point_mask(v).cpu().numpy()
        n, rmu, rS = _ref_moments(x.cpu().numpy(), m)
        assert cnt[k].item() == n
        assert np.allclose(mu[k].numpy(), rmu, rtol=0, atol=1e-12 * 700)
        assert np.abs(S[k].numpy() - rS).max() <= 1e-11 * np.abs(rS).max()
    # a fixed summation order: the same bits again
    cnt2, mu2, S2 = zag._moments(ctx, xs, sels)
    assert torch.equal(mu, mu2) and torch.equal(S, S2) and torch.equal(cnt, cnt2)


def _torch_loss(x, m):
    """The residual-sum loss in torch (FP64, host): n from the scatter's
    smallest eigenvector, held fixed; the mean differentiable."""
    p = torch.where(m[:, None], x, 0.0)
    mu = p.sum(0) / m.sum()
    d = p - mu
    with torch.no_grad():
        dd = torch.where(m[:, None], d, 0.0)
        n = torch.linalg.eigh(dd.T @ dd)[1][:, 0]
    r = torch.where(m, d @ n, 0.0)
    return (r * r).sum()


@gpu
def test_plane_losses_and_gradient_match_torch(ctx):
    views = _views(1)
    xs = [x.clone().requires_grad_(True) for x, _ in views]
    loss = zag.plane_losses([(x, v) for x, (_, v) in zip(xs, views)], ctx)
    w = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64)
    (loss * w).sum().backward()
    for k, (x, v) in enumerate(views):
        xr = x.detach().cpu().clone().requires_grad_(True)
        m = zag.point_mask(v).cpu()
        lr = _torch_loss(xr, m)
        (lr * w[k]).backward()
        assert abs(loss[k].item() - lr.item()) <= 1e-9 * lr.item()
        gd, gr = xs[k].grad.cpu(), xr.grad
        assert (gd - gr).abs().max().item() <= 1e-9 * gr.abs().max().item()
        assert torch.equal(gd[~m], torch.zeros_like(gd[~m]))


@gpu
def test_plane_flatness_gradient_matches_torch(ctx):
    views = _views(2)
    xs = [x.clone().requires_grad_(True) for x, _ in views]
    fl = zag.plane_flatnesses([(x, v) for x, (_, v) in zip(xs, views)], ctx)
    fl.sum().backward()
    for k, (x, v) in enumerate(views):
        xr = x.detach().cpu().clone().requires_grad_(True)
        m = zag.point_mask(v).cpu()
        p = torch.where(m[:, None], xr, 0.0)
        mu = p.sum(0) / m.sum()
        d = torch.where(m[:, None], p - mu, 0.0)
        ev = torch.linalg.eigvalsh(d.T @ d)
        f = ev[0] / ev.sum()
        f.backward()
        assert abs(fl[k].item() - f.item()) <= 1e-9 * abs(f.item())
        gd, gr = xs[k].grad.cpu(), xr.grad
        assert (gd - gr).abs().max().item() <= 1e-8 * gr.abs().max().item()


@gpu
def test_empty_selection(ctx):
    x = torch.zeros((10, 3), dtype=torch.float64, device="cuda")
    sel = torch.zeros(10, dtype=torch.bool, device="cuda")
    cnt, mu, S = zag._moments(ctx, [x], [zag._sel(sel)])
    assert cnt[0].item() == 0 and not mu.any() and not S.any()


@gpu
def test_plane_loss_on_another_context_sees_the_triangulation():
    """Points triangulated through one context and scored through another
    (here the process-wide default): with torch on its default stream both
    contexts run on the legacy null stream, so the second sees the first's
    results. Every evaluation equals the same-context loss bit for bit."""
    import zim_amd
    from tests.test_gpu_adjoint import perturbed, plane_pair
    ctx, phL, phR, mc, theta0, nominal = plane_pair(256, 320)
    other = zim_amd.default_context()
    assert other is not ctx
    for k in range(4):
        th = perturbed(theta0, 20 + k)
        with torch.no_grad():
            xyz, ok = zag.triangulate(th, phL, phR, mc, nominal, ctx)
            a = zag.plane_loss(xyz, ok, other).item()
            xyz2, ok2 = zag.triangulate(th, phL, phR, mc, nominal, ctx)
            torch.cuda.synchronize()
            b = zag.plane_loss(xyz2, ok2, ctx).item()
        assert a == b, (k, a, b)
