"""The CPU restatement of the unorganized-cloud filters (voxel grid) pinned
by brute force and closed forms: exact k nearest distances, the radius
neighbour sets, and the organized MLS filter on a regular grid where the
pixel window and the radius pick the same neighbours."""
import numpy as np
import pytest

import oracle
from tests import scenes


def _rng_cloud(n, seed, holes=0.0, spread=4.0):
    g = np.random.default_rng(seed)
    xy = g.uniform(0, spread, (n, 2))
    z = 700 + 0.3 * np.sin(xy[:, 0]) * np.cos(0.7 * xy[:, 1]) + g.normal(0, 0.02, n)
    xyz = np.column_stack([xy, z]).astype(np.float32)
    valid = (g.random(n) >= holes).astype(np.uint8)
    return xyz, valid


@pytest.mark.parametrize("k,voxel,rmax", [(8, 0.2, 4), (30, 0.15, 4), (30, 0.12, 3), (1, 1.0, 1)])
def test_cloud_sor_is_exact_knn(k, voxel, rmax):
    """Eligible points are exactly those whose k-th nearest lies within rmax
    voxels, and their mean distance is the mean of the brute-force k nearest
    distances, summed ascending: bit-exact."""
    xyz, valid = _rng_cloud(1500, 3, holes=0.1)
    keep, md, st = oracle.cloud_sor(xyz, valid, k, 1.0, voxel, rmax)
    d2, _ = oracle.cloud_brute(xyz, valid, k, 0.0)
    lim = np.float32(voxel * rmax) ** 2 * np.float32(0.99999)
    eligible = (valid != 0) & (d2[:, k - 1] <= lim)
    assert eligible.sum() > 100
    ref = np.zeros(len(xyz))
    for i in np.nonzero(eligible)[0]:
        s = 0.0
        for v in d2[i]:
            s += np.sqrt(np.float64(v))
        ref[i] = s / k
    assert np.array_equal(md, ref)
    mean = ref[eligible].mean()
    sd = ref[eligible].std(ddof=1)
    assert st[0] == eligible.sum()
    assert abs(st[1] - mean) < 1e-12 and abs(st[2] - sd) < 1e-12
    assert np.array_equal(keep != 0, eligible & (ref <= st[3]))


@pytest.mark.parametrize("voxel", [0.45, 0.21])   # 3x3x3 cells; 5x5x5 half-radius cells
def test_cloud_gaussian_brute_force(voxel):
    xyz, valid = _rng_cloud(1200, 5, holes=0.05)
    sigma, radius = 0.15, 0.4
    got = oracle.cloud_gaussian(xyz, valid, sigma, radius, voxel)
    x = xyz.astype(np.float64)
    ok = valid != 0
    for i in range(0, len(xyz), 7):
        if not ok[i]:
            assert np.array_equal(got[i], xyz[i])
            continue
        d2 = ((x - x[i]) ** 2).sum(1)
        m = ok & (d2 <= radius * radius * (1 - 1e-6))
        w = np.exp(-d2[m] / (2 * sigma * sigma))
        ref = (w[:, None] * x[m]).sum(0) / w.sum()
        # FP32 weighted sums of ~30 terms around z = 700 mm: 1e-6 relative
        assert np.abs(got[i] - ref).max() < 1e-6 * 700


@pytest.mark.parametrize("voxel", [3.0, 1.5])
def test_cloud_mls_matches_organized_window(voxel):
    """Points on a regular unit grid in x, y with a gentle height: a 5 x 5
    pixel window with a 2.9 cutoff and a 2.9 radius select the same
    neighbours, so the organized filter (the reference's) and the cloud
    filter agree up to the order of summation. Order 1: on a lattice the two
    in-plane eigenvalues nearly coincide, and an order-2 fit amplifies the
    reordering far more (the quadric test below pins order 2)."""
    order = 1
    rows, cols = 40, 48
    ii, jj = np.mgrid[0:rows, 0:cols].astype(np.float64)
    z = 0.4 * np.sin(jj / 6.0) * np.cos(ii / 9.0)
    org = np.stack([jj, ii, z], -1).astype(np.float32)
    P = np.ones((rows, cols), np.uint8)
    ref = oracle.mls(org, P, 2, 1.2, order, 2.9, 1, 1)
    got = oracle.cloud_mls(org.reshape(-1, 3), None, 1.2, order, 2.9, 1, 1, voxel).reshape(rows, cols, 3)
    d = np.abs(got - ref).max(-1)
    # the corner points (a quarter disc of neighbours) are the least well
    # conditioned: there the summation order alone moves the fit by ~2e-3
    assert d[2:-2, 2:-2].max() < 5e-4 and d.max() < 5e-3 and np.median(d) < 1e-6
    assert np.abs(got - org).max() > 1e-3   # the filter moved points


@pytest.mark.parametrize("voxel", [1.2, 0.6])
def test_cloud_mls_fixed_points_and_iterations(voxel):
    """A plane is a fixed point (order 1); a quadric height field is
    reproduced by order 2 far better than by order 1 (not exactly: in the
    tilted local frame the surface picks up cubic terms); invalid points pass
    through unchanged."""
    g = np.random.default_rng(2)
    xy = g.uniform(-3, 3, (2000, 2))
    plane = np.column_stack([xy, 10 + 0.2 * xy[:, 0] - 0.1 * xy[:, 1]]).astype(np.float32)
    out = oracle.cloud_mls(plane, None, 0.5, 1, 1.2, 2, 1, voxel)
    assert np.abs(out - plane).max() < 2e-5
    quad = np.column_stack([xy, 0.05 * xy[:, 0] ** 2 - 0.03 * xy[:, 0] * xy[:, 1]
                            + 0.02 * xy[:, 1] ** 2]).astype(np.float32)
    valid = (g.random(2000) > 0.1).astype(np.uint8)
    out = oracle.cloud_mls(quad, valid, 0.5, 2, 1.2, 1, 1, voxel)
    out1 = oracle.cloud_mls(quad, valid, 0.5, 1, 1.2, 1, 1, voxel)
    e2, e1 = np.abs(out - quad).max(), np.abs(out1 - quad).max()
    assert e2 < 5e-4 and e2 < 0.2 * e1, (e2, e1)
    assert np.array_equal(out[valid == 0], quad[valid == 0])


def test_cloud_voxel_below_radius_is_rejected():
    """Radius filters need voxel >= radius / 2 (5x5x5 cells cover the ball)."""
    xyz, _ = _rng_cloud(50, 1)
    oracle.cloud_gaussian(xyz, None, 0.1, 0.3, 0.15)
    with pytest.raises(AssertionError):
        oracle.cloud_gaussian(xyz, None, 0.1, 0.3, 0.14)


def _exact_cells(v, o, voxel):
    """floor((v - o) / voxel) in exact rational arithmetic, and whether the
    quotient sits within 1e-9 cells of an integer (a tie the pin skips)."""
    from fractions import Fraction
    fo, fv = Fraction(float(np.float32(o))), Fraction(float(np.float32(voxel)))
    cells, near = [], []
    for x in np.asarray(v, np.float32):
        q = (Fraction(float(x)) - fo) / fv
        c = q.numerator // q.denominator
        cells.append(c)
        near.append(min(q - c, c + 1 - q) < Fraction(1, 10 ** 9))
    return np.array(cells, np.int64), np.array(near)


@pytest.mark.parametrize("voxel", [0.3, 0.7, 1.1])
def test_voxel_cells_are_exact_across_a_wide_grid(voxel):
    """Cell assignment of the shared voxel frame across about 2000 cells: the
    exact floor((v - o) / voxel) for random coordinates and for coordinates
    within a few ulps of cell boundaries. The same coordinates assigned in
    float, (v - o) * (1 / voxel) rounded at each step, miss by a cell on
    some of them: the data would catch that."""
    g = np.random.default_rng(int(voxel * 10))
    o = np.float32(-2.0) * np.float32(voxel)
    v = g.uniform(0, 2000 * voxel, 4000).astype(np.float32)
    bnd = (o.astype(np.float64) + voxel * np.arange(900, 2000, 7)).astype(np.float32)
    steps = np.arange(-6, 7, dtype=np.float32)
    v = np.concatenate([v, (bnd[:, None] + steps[None, :] * np.spacing(bnd)[:, None]).ravel()])
    got = oracle.voxel_cells(v, o, voxel)
    want, near = _exact_cells(v, o, voxel)
    assert near.sum() < 5
    assert np.array_equal(got[~near], want[~near])
    in_float = np.floor((v - o) * (np.float32(1) / np.float32(voxel))).astype(np.int64)
    assert (in_float[~near] != want[~near]).sum() > 0


def test_knn_point_limit_bounds_every_point_outside_the_block():
    """The device's first k-nearest pass settles a point whose k-th squared
    distance within its 3x3x3 block of cells is at most the point's own
    bound: one voxel plus its distance to the nearest face of its cell, less
    a margin. Every point outside the block must then be farther. Points
    placed on and just past the block's faces (kept when the exact cell
    assignment puts them outside), straight across from the query or offset
    along another axis, all have float squared distances (as the kernel
    computes them) above the bound, and the closest come within 1e-4 of it,
    so the bound is the face distance and no looser; the same bound taken to
    the farthest face of the cell is exceeded by some of them."""
    g = np.random.default_rng(5)
    voxel = np.float32(0.7)
    o = np.array([-3.1, 12.5, 480.0], np.float32)
    n = 20000
    q = (o + g.uniform(2, 900, (n, 3)) * voxel).astype(np.float32)
    lim = oracle.knn_point_limit(q, o, voxel)
    assert (lim >= voxel * voxel * np.float32(0.9999)).all()
    assert (lim <= (np.float32(1.5) * voxel) ** 2).all()
    # planted: the bound to the farthest face of the point's own cell
    t = (q.astype(np.float64) - o) / np.float64(voxel)
    t -= np.floor(t)
    far = (np.float64(voxel) * (1 + np.minimum(t, 1 - t).max(axis=1))) ** 2 * (1 - 4e-5)
    qc = np.stack([oracle.voxel_cells(q[:, c], o[c], voxel) for c in range(3)], 1)
    d2s, idx = [], []
    for a in range(3):
        for side in (-1, 1):
            # the nearest face of the neighbouring block cell, a few ulps on
            # either side
            face = (o[a].astype(np.float64) +
                    (qc[:, a] + (2 if side > 0 else -1)) * np.float64(voxel)).astype(np.float32)
            for ulps in (-2, -1, 0, 1, 2):
                for off in (0.0, 0.4):
                    p = q.copy()
                    p[:, a] = face + np.float32(ulps) * np.spacing(face)
                    if off:
                        b = (a + 1) % 3
                        p[:, b] = (p[:, b] + g.uniform(-off, off, n) * voxel).astype(np.float32)
                    pc = np.stack([oracle.voxel_cells(p[:, c], o[c], voxel) for c in range(3)], 1)
                    out = np.flatnonzero((np.abs(pc - qc) >= 2).any(axis=1))
                    d = (q[out] - p[out]).astype(np.float32)
                    d2s.append((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
                    idx.append(out)
    d2, idx = np.concatenate(d2s), np.concatenate(idx)
    assert len(d2) > 200000
    assert (d2 > lim[idx]).all()
    assert (d2 / lim[idx]).min() < 1 + 1e-4
    assert (d2 <= far[idx]).any()


def test_frame_rejects_more_than_2_30_cells_along_an_axis():
    """Cell assignment stays within 2^-22 cells of exact only up to 2^30 cells
    along an axis; a frame longer than that is refused even when its total
    cell count is allowed."""
    ok = np.array([[0, 0, 0], [1.0e6, 0, 0]], np.float32)        # 1.0e9 cells along x
    oracle.cloud_sor(ok, None, 1, 1.0, 1e-3)
    with pytest.raises(AssertionError):
        oracle.cloud_sor(ok * np.float32(1.1), None, 1, 1.0, 1e-3)  # 1.1e9


def test_filters_exact_on_a_cloud_a_thousand_voxels_wide():
    """kNN (SOR) and radius (Gaussian) neighbour sets on a cloud spanning
    about 1000 voxels, by brute force: strips of points across the whole
    extent, so neighbours meet at cell boundaries far from the origin."""
    g = np.random.default_rng(11)
    voxel, radius = 0.3, 0.3
    xs = g.uniform(0, 1000 * voxel, 3000)
    ys = g.uniform(0, 1.5, 3000)
    zs = 500 + g.normal(0, 0.05, 3000)
    xyz = np.column_stack([xs, ys, zs]).astype(np.float32)
    k = 8
    keep, md, st = oracle.cloud_sor(xyz, None, k, 1.0, voxel, 4)
    d2, within = oracle.cloud_brute(xyz, None, k, radius)
    lim = np.float32(voxel * 4) ** 2 * np.float32(0.99999)
    eligible = d2[:, k - 1] <= lim
    assert eligible.sum() > 1000
    ref = np.zeros(len(xyz))
    for i in np.nonzero(eligible)[0]:
        s = 0.0
        for v in d2[i]:
            s += np.sqrt(np.float64(v))
        ref[i] = s / k
    assert np.array_equal(md, ref)
    # Gaussian: the weighted mean over exactly the brute-force radius set
    got = oracle.cloud_gaussian(xyz, None, 0.15, radius, voxel)
    x = xyz.astype(np.float64)
    r2 = np.float32(radius) * np.float32(radius)
    for i in range(0, len(xyz), 3):
        dd = ((xyz - xyz[i]) ** 2).astype(np.float32)
        m = (dd[:, 0] + dd[:, 1] + dd[:, 2]) <= r2
        assert m.sum() == within[i]
        w = np.exp(-((x[m] - x[i]) ** 2).sum(1) / (2 * 0.15 * 0.15))
        ref_i = (w[:, None] * x[m]).sum(0) / w.sum()
        assert np.abs(got[i] - ref_i).max() < 1e-6 * 500
