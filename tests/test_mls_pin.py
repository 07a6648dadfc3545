"""Pins the oracle's MLS filter (JSLAB_MLS_Filter.cpp:11-441, GPU fallback
JSLAB_MLS_Filter.cu:283-286) against an independent FP64 statement of
weighted moving least squares, written here from its textbook definition:

  neighbours   the valid points of the (2hw+1)^2 window with |p - q|^2 <= r^2
               (r = max_radius, or 3 sigma when it is 0), the point itself
               included;
  weights      w = exp(-|p - q|^2 / (2 sigma^2));
  frame        weighted centroid c and covariance; n = the eigenvector of its
               smallest eigenvalue (numpy.linalg.eigh), t1, t2 the others;
  fit          weighted least squares (numpy.linalg.lstsq on sqrt(w) rows) of
               f = (q - c).n over the basis {1, u, v[, u^2, uv, v^2]} with
               u = (q - c).t1, v = (q - c).t2;
  projection   c + u0 t1 + v0 t2 + f(u0, v0) n at the point's own u0, v0.

The result of orders 1 and 2 does not depend on which tangent basis eigh
returns (the basis spans all polynomials of that degree in u, v) nor on the
sign of n, so the statement needs none of the reference's eigen-solver
conventions. Noise and a radius smaller than the window make the weights and
the cutoff matter; the planted-mistake tests show that a misread sigma, a
weight without the factor 2, or a cutoff 10 % off would fail the bar.

The fallback ladder is checked against a table written out by hand from the
reference (CPU: JSLAB_MLS_Filter.cpp:273-275; GPU: JSLAB_MLS_Filter.cu:283-286),
and the solver's pieces (eigen3, Cholesky) against numpy.linalg.
"""
import math

import numpy as np
import pytest

import oracle

# ---- the FP64 statement --------------------------------------------------------

# Effective polynomial order by (semantics, requested order) for neighbour
# counts n = 0..10; None = fewer than 3 neighbours, the point is left as it is.
# CPU (JSLAB_MLS_Filter.cpp:273-275): order if n >= {3,5,8}[order], else 0,
# then a 1/0 fallback that can no longer trigger. GPU (.cu:283-286): the
# requested order, lowered to 1 (n >= 5) or 0 when n < {3,5,8}[order].
_N = None
LADDER = {
    # n:          0   1   2   3  4  5  6  7  8  9  10
    (0, 0): [_N, _N, _N, 0, 0, 0, 0, 0, 0, 0, 0],
    (0, 1): [_N, _N, _N, 0, 0, 1, 1, 1, 1, 1, 1],
    (0, 2): [_N, _N, _N, 0, 0, 0, 0, 0, 2, 2, 2],
    (1, 0): [_N, _N, _N, 0, 0, 0, 0, 0, 0, 0, 0],
    (1, 1): [_N, _N, _N, 0, 0, 1, 1, 1, 1, 1, 1],
    (1, 2): [_N, _N, _N, 0, 0, 1, 1, 1, 2, 2, 2],
}


def textbook_point(p, nbrs, sigma, r, order):
    """Projection of p (f64[3]) onto the order-`order` weighted MLS surface of
    `nbrs` (f64[m,3], every candidate; the radius test is applied here).
    Returns (projected point, neighbour count, condition number of the normal
    equations, distance of the nearest d^2 to the cutoff)."""
    d2 = ((nbrs - p) ** 2).sum(1)
    inr = d2 <= r * r
    q = nbrs[inr]
    w = np.exp(-d2[inr] / (2.0 * sigma * sigma))
    c = (w[:, None] * q).sum(0) / w.sum()
    e = q - c
    cov = np.einsum("k,ki,kj->ij", w, e, e)
    _, V = np.linalg.eigh(cov)
    n, t1, t2 = V[:, 0], V[:, 1], V[:, 2]
    e0 = p - c
    margin = np.abs(d2 - r * r).min()
    if order == 0:
        return p - (e0 @ n) * n, len(q), 1.0, margin
    u, v, f = e @ t1, e @ t2, e @ n

    def basis(a, b):
        cols = [np.ones_like(a), a, b]
        if order == 2:
            cols += [a * a, a * b, b * b]
        return np.stack(cols, -1)

    A = basis(u, v)
    sw = np.sqrt(w)
    co = np.linalg.lstsq(A * sw[:, None], f * sw, rcond=None)[0]
    u0, v0 = e0 @ t1, e0 @ t2
    ff = basis(np.array(u0), np.array(v0)) @ co
    cond = np.linalg.cond((A * w[:, None]).T @ A)
    return c + u0 * t1 + v0 * t2 + ff * n, len(q), cond, margin


def textbook_organized(xyz, P, hw, sigma, order, max_radius, iterations, sem):
    """The filter over an organized cloud; returns (out f64, per-point info:
    neighbour count, condition, cutoff margin) of the last iteration."""
    rows, cols = P.shape
    r = max_radius if max_radius > 0 else 3.0 * sigma
    cur = xyz.astype(np.float64)
    for _ in range(max(1, iterations)):
        out = cur.copy()
        info = np.zeros((rows, cols, 3))
        for i in range(rows):
            for j in range(cols):
                if not P[i, j]:
                    continue
                r0, r1 = max(0, i - hw), min(rows, i + hw + 1)
                c0, c1 = max(0, j - hw), min(cols, j + hw + 1)
                win = cur[r0:r1, c0:c1][P[r0:r1, c0:c1] > 0]
                nv = int((((win - cur[i, j]) ** 2).sum(1) <= r * r).sum())
                eff = LADDER[(sem, order)][min(nv, 10)]
                if eff is None:
                    info[i, j] = (nv, 1.0, 1.0)
                    continue
                out[i, j], *inf = textbook_point(cur[i, j], win, sigma, r, eff)
                info[i, j] = inf
        cur = out
    return cur, info


def noisy_patch(rows, cols, pitch, z0, noise, seed, holes=0.08):
    """A tilted curved surface sampled on a jittered grid with Gaussian noise
    along its normal and random holes (P = 0)."""
    g = np.random.default_rng(seed)
    u, v = np.meshgrid(np.arange(cols) * pitch, np.arange(rows) * pitch)
    u = u + g.uniform(-0.05, 0.05, u.shape) * pitch
    v = v + g.uniform(-0.05, 0.05, v.shape) * pitch
    a, b = u - u.mean(), v - v.mean()
    w = 0.03 * a * a - 0.02 * a * b + 0.025 * b * b + 0.1 * np.sin(a) + g.normal(0, noise, a.shape)
    th = 0.4
    x = u
    y = v * math.cos(th) - w * math.sin(th)
    z = z0 + v * math.sin(th) + w * math.cos(th)
    P = (g.random((rows, cols)) > holes).astype(np.uint8)
    return np.stack([x, y, z], -1).astype(np.float32), P


# FP32 coordinates: the bar is a small multiple of the coordinates' ulp
# (9.5e-7 near 8 mm, 6.1e-5 near 700 mm). Near 700 mm a plane projection
# (order 0) also carries the tilt of the FP32 normal: the covariance is built
# from offsets rounded at 3e-5 mm against an eigenvalue gap of ~0.1 mm^2, a
# tilt of ~1e-4 rad, times the point's offset from the centroid along the
# surface (up to 1 mm at the patch's edges). Orders 1 and 2 absorb a tilt
# into their linear terms.
BARS = {(5.0, 0): 1e-5, (5.0, 1): 1e-5, (5.0, 2): 1e-5,
        (700.0, 0): 3e-4, (700.0, 1): 1e-4, (700.0, 2): 1e-4}


def _well_conditioned(P, info):
    """Points whose fit the FP32 filter can reproduce: order-2 normal equations
    with condition < 1e4, and no neighbour within 1e-4 mm^2 of the cutoff
    (where FP32 and FP64 distances may fall on different sides)."""
    return (P > 0) & (info[..., 1] < 1e4) & (info[..., 2] > 1e-4)


@pytest.mark.parametrize("z0", [5.0, 700.0])
@pytest.mark.parametrize("sem", [0, 1])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_oracle_mls_equals_fp64_textbook(z0, sem, order):
    xyz, P = noisy_patch(22, 24, 0.3, z0, 0.03, seed=1)
    hw, sigma, r = 5, 0.5, 1.0           # the 11 x 11 window reaches 1.5 mm > r
    ref, info = textbook_organized(xyz, P, hw, sigma, order, r, 1, sem)
    got = oracle.mls(xyz, P, hw, sigma, order, r, 1, sem)
    ok = _well_conditioned(P, info)
    assert ok.sum() >= 0.9 * (P > 0).sum()
    d = np.abs(got - ref).max(-1)
    assert d[ok].max() < BARS[z0, order], d[ok].max()
    # the filter did move the points (by about the noise), so the bar means something
    assert np.abs(ref - xyz).max(-1)[ok].mean() > 50 * BARS[z0, order]
    assert np.array_equal(got[P == 0], xyz[P == 0])


@pytest.mark.parametrize("sem", [0, 1])
def test_oracle_mls_two_iterations_equal_textbook(sem):
    xyz, P = noisy_patch(20, 22, 0.3, 5.0, 0.03, seed=4)
    ref, info = textbook_organized(xyz, P, 5, 0.5, 2, 1.0, 2, sem)
    got = oracle.mls(xyz, P, 5, 0.5, 2, 1.0, 2, sem)
    ok = _well_conditioned(P, info)
    assert ok.sum() >= 0.9 * (P > 0).sum()
    # the second pass starts from the first pass's FP32 rounding: twice the bar
    assert np.abs(got - ref).max(-1)[ok].max() < 2e-5


def test_oracle_mls_window_clips_radius_and_default_cutoff():
    """A 5 x 5 window (0.6 mm at 0.3 mm pitch) holds fewer points than the
    radius; max_radius = 0 selects the 3 sigma cutoff."""
    xyz, P = noisy_patch(20, 22, 0.3, 5.0, 0.03, seed=5)
    for hw, r_arg in ((2, 1.0), (5, 0.0)):
        ref, info = textbook_organized(xyz, P, hw, 0.3, 2, r_arg, 1, 1)
        got = oracle.mls(xyz, P, hw, 0.3, 2, r_arg, 1, 1)
        ok = _well_conditioned(P, info)
        assert ok.sum() >= 0.8 * (P > 0).sum()
        assert np.abs(got - ref).max(-1)[ok].max() < 1e-5, (hw, r_arg)


@pytest.mark.parametrize("mistake", ["sigma*1.1", "no factor 2", "radius*0.9", "radius*1.1",
                                     "window too small"])
def test_planted_mistakes_fail_the_bar(mistake):
    """The oracle run with a misread parameter misses the textbook at the
    bar used above by a wide margin, on the same points."""
    xyz, P = noisy_patch(22, 24, 0.3, 5.0, 0.03, seed=1)
    hw, sigma, r = 5, 0.5, 1.0
    ref, info = textbook_organized(xyz, P, hw, sigma, 2, r, 1, 1)
    kw = {"hw": hw, "sigma": sigma, "r": r}
    if mistake == "sigma*1.1":
        kw["sigma"] = sigma * 1.1
    elif mistake == "no factor 2":      # w = exp(-d^2 / sigma^2)
        kw["sigma"] = sigma / math.sqrt(2.0)
    elif mistake == "radius*0.9":
        kw["r"] = r * 0.9
    elif mistake == "radius*1.1":
        kw["r"] = r * 1.1
    else:
        # a 5 x 5 window at 0.3 mm pitch no longer holds the 1 mm radius
        kw["hw"] = 2
    got = oracle.mls(xyz, P, kw["hw"], kw["sigma"], 2, kw["r"], 1, 1)
    ok = _well_conditioned(P, info)
    assert np.abs(got - ref).max(-1)[ok].max() > 100 * 1e-5


def test_cloud_mls_equals_fp64_textbook():
    """The unorganized-cloud MLS (radius neighbours from the voxel grid)
    against the textbook over brute-force radius neighbours."""
    g = np.random.default_rng(7)
    n = 1500
    xy = g.uniform(0, 6, (n, 2))
    a, b = xy[:, 0] - 3, xy[:, 1] - 3
    z = 4 + 0.04 * a * a - 0.03 * a * b + 0.02 * b * b + 0.1 * np.sin(a) + g.normal(0, 0.03, n)
    th = 0.3
    pts = np.column_stack([xy[:, 0], xy[:, 1] * math.cos(th) - z * math.sin(th),
                           xy[:, 1] * math.sin(th) + z * math.cos(th)]).astype(np.float32)
    valid = (g.random(n) > 0.05).astype(np.uint8)
    sigma, r = 0.4, 0.8
    for sem in (0, 1):
        got = oracle.cloud_mls(pts, valid, sigma, 2, r, 1, sem, 0.8)
        X = pts.astype(np.float64)
        V = X[valid > 0]
        checked = 0
        for i in range(n):
            if not valid[i]:
                assert np.array_equal(got[i], pts[i])
                continue
            nv = int((((V - X[i]) ** 2).sum(1) <= r * r).sum())
            eff = LADDER[(sem, 2)][min(nv, 10)]
            if eff is None:
                assert np.array_equal(got[i], pts[i])
                continue
            ref, _, cond, margin = textbook_point(X[i], V, sigma, r, eff)
            if cond < 1e4 and margin > 1e-4:
                assert np.abs(got[i] - ref).max() < 1e-5, (i, sem)
                checked += 1
        assert checked > 0.8 * valid.sum()


def _ladder_window(n, seed):
    """A 7 x 7 window whose centre has exactly n valid neighbours (itself
    included), all well inside the radius, on a noisy curved surface."""
    g = np.random.default_rng(seed)
    xyz, _ = noisy_patch(7, 7, 0.3, 5.0, 0.05, seed=seed, holes=0.0)
    P = np.zeros((7, 7), np.uint8)
    P[3, 3] = 1
    others = [(i, j) for i in range(7) for j in range(7) if (i, j) != (3, 3)]
    for k in g.permutation(len(others))[: n - 1]:
        P[others[k]] = 1
    return xyz, P


@pytest.mark.parametrize("sem", [0, 1])
def test_fallback_ladder_matches_the_hand_table(sem):
    """For neighbour counts 1..10 and each requested order, the oracle's
    output at the centre equals, bit for bit, its output at the order the
    table names; and the three orders give different bits, so the check can
    tell them apart."""
    hw, sigma, r = 3, 0.6, 3.0
    distinct = 0
    for n in range(1, 11):
        for seed in range(3):
            xyz, P = _ladder_window(n, 100 * n + seed)
            outs = {o: oracle.mls(xyz, P, hw, sigma, o, r, 1, sem)[3, 3] for o in (0, 1, 2)}
            for req in (0, 1, 2):
                eff = LADDER[(sem, req)][n]
                if eff is None:
                    assert np.array_equal(outs[req], xyz[3, 3]), (n, req)
                    continue
                # the effective order, forced: request it where the table maps
                # it to itself (order 0 always; 1 and 2 from 5 and 8 on)
                assert LADDER[(sem, eff)][n] == eff
                assert np.array_equal(outs[req], outs[eff]), (sem, n, req, eff)
            if n >= 8:
                distinct += len({tuple(v) for v in outs.values()}) == 3
            if 5 <= n < 8:
                distinct += not np.array_equal(outs[0], outs[1])
    assert distinct >= 12


def test_eigen3_matches_eigh():
    g = np.random.default_rng(3)
    mats = []
    for _ in range(400):
        Q, _ = np.linalg.qr(g.normal(size=(3, 3)))
        lam = np.sort(g.uniform(0.01, 1.0, 3)) * 10.0 ** g.uniform(-3, 3)
        mats.append((Q * lam) @ Q.T)
    C = np.array(mats, np.float32)
    C = (C + C.transpose(0, 2, 1)) / 2
    ev, V = oracle.eigen3(C)
    lam, W = np.linalg.eigh(C.astype(np.float64))
    scale = np.abs(lam).max(-1, keepdims=True)
    assert (np.abs(ev - lam) / scale).max() < 2e-5
    # eigenvectors (one per row) up to sign where the eigenvalue is separated
    for i in range(len(C)):
        gaps = [min(abs(lam[i, j] - lam[i, k]) for k in range(3) if k != j) / scale[i, 0]
                for j in range(3)]
        for j in range(3):
            if gaps[j] > 1e-2:
                assert abs(V[i, j] @ W[i][:, j]) > 1 - 1e-4, (i, j)
    # diagonal input (off-diagonals below 1e-20): sorted diagonal, unit vectors
    ev, V = oracle.eigen3(np.diag([3.0, 1.0, 2.0]).astype(np.float32))
    assert list(ev) == [1.0, 2.0, 3.0]
    assert np.array_equal(np.abs(V), np.eye(3)[[1, 2, 0]])


def test_cholesky_matches_solve_and_rejects_small_pivots():
    g = np.random.default_rng(9)
    for nc in range(1, 7):
        for _ in range(50):
            M = g.normal(size=(nc + 3, nc))
            A = (M.T @ M).astype(np.float32)
            b = g.normal(size=nc).astype(np.float32)
            x, ok = oracle.chol(A, b)
            assert ok
            ref = np.linalg.solve(A.astype(np.float64), b.astype(np.float64))
            cond = np.linalg.cond(A.astype(np.float64))
            assert np.abs(x - ref).max() <= 1e-6 * cond * max(1.0, np.abs(ref).max())
    # rank-deficient and indefinite systems fail; so does a pivot of 1e-12
    v = g.normal(size=(4, 1))
    assert not oracle.chol((v @ v.T).astype(np.float32), np.ones(4, np.float32))[1]
    assert not oracle.chol(np.diag([1.0, -1.0, 1.0]).astype(np.float32), np.ones(3, np.float32))[1]
    assert not oracle.chol(np.diag([1.0, np.float32(1e-12)]).astype(np.float32),
                           np.ones(2, np.float32))[1]
    assert oracle.chol(np.diag([1.0, 2e-12]).astype(np.float32), np.ones(2, np.float32))[1]
