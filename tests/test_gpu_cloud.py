"""Unorganized-cloud filters on the device (voxel grid, LDS neighbour
tiles) against the CPU restatement: SOR masks and mean distances
bit-exact, MLS and Gaussian within 1e-3 mm (bit-identical in practice: the
same candidate order and arithmetic), on random clouds, on merged clouds
(noisy views of one C2 scene; ring views of the plane + spheres and boxes
scene), and on edge cases."""
import numpy as np
import pytest
import torch

import oracle
import zim_amd
from tests import scenes
from tests.test_cloud_cpu import _rng_cloud

pytestmark = pytest.mark.gpu

XYZ_TOL = 1e-3


@pytest.fixture(scope="module")
def ctx():
    return zim_amd.Context(0)


def _sor_both(ctx, xyz, valid, k, voxel, rmax, dev=True):
    ref = oracle.cloud_sor(xyz, valid, k, 1.0, voxel, rmax)
    if dev:
        got = zim_amd.cloud_sor(torch.from_numpy(xyz).cuda(),
                                None if valid is None else torch.from_numpy(valid).cuda(),
                                k, 1.0, voxel, rmax, ctx)
        torch.cuda.synchronize()
        got = [g.cpu().numpy() for g in got]
    else:
        got = zim_amd.cloud_sor(xyz, valid, k, 1.0, voxel, rmax, ctx)
    return ref, got


def _assert_sor(ref, got):
    assert np.array_equal(got[1], ref[1]), np.abs(got[1] - ref[1]).max()
    assert np.array_equal(got[0], ref[0])
    assert np.array_equal(got[2], ref[2])


# the mean points per occupied cell selects the kNN and Gaussian workgroup
# width: voxels of 0.2-0.45 give one wave, 0.6 two, 1.0 four on these clouds
@pytest.mark.parametrize("k,voxel,rmax", [(8, 0.2, 4), (30, 0.15, 4), (30, 0.12, 3), (20, 0.3, 1),
                                          (50, 0.2, 4), (13, 0.25, 2), (30, 0.6, 2), (50, 1.0, 2)])
def test_cloud_sor_random(ctx, k, voxel, rmax):
    xyz, valid = _rng_cloud(6000, 11, holes=0.1)
    ref, got = _sor_both(ctx, xyz, valid, k, voxel, rmax)
    _assert_sor(ref, got)
    assert ref[0].sum() > 100


def test_cloud_sor_retry_waves_near_and_far(ctx):
    """The second k-nearest pass takes the unsettled points 64 at a time in
    grid order and walks the union of their search boxes when those lie close
    together, each point's own box otherwise. A dense patch (its unsettled
    points neighbours in the grid: compact unions) beside sparse points
    strung along y and z (consecutive in grid order but rows of cells apart:
    boxes walked one by one) give both; the k nearest, the eligibility and
    the statistics equal the oracle's bit for bit."""
    g = np.random.default_rng(21)
    patch = np.column_stack([g.uniform(0, 2.0, 3000), g.uniform(0, 2.0, 3000),
                             500 + g.normal(0, 0.05, 3000)])
    t = np.sort(g.uniform(0, 40.0, 600))
    strand = np.column_stack([5.0 + g.normal(0, 0.3, 600), t, 500 + 0.5 * t + g.normal(0, 0.3, 600)])
    xyz = np.concatenate([patch, strand]).astype(np.float32)
    for k, voxel, rmax in ((20, 0.12, 4), (30, 0.15, 6), (50, 0.2, 8)):
        ref, got = _sor_both(ctx, xyz, None, k, voxel, rmax)
        _assert_sor(ref, got)
        assert 0 < ref[0].sum() < len(xyz)


def test_cloud_sor_host_mode_and_edge_cases(ctx):
    xyz, valid = _rng_cloud(3000, 12)
    xyz[5] = np.nan                      # not finite: removed
    xyz[7] = xyz[8]                      # a duplicate: distance 0
    xyz[9:20] = xyz[9]                   # a cluster of 11 identical points
    ref, got = _sor_both(ctx, xyz, None, 10, 0.2, 4, dev=False)
    _assert_sor(ref, got)
    assert got[0][5] == 0 and got[1][5] == 0.0
    # nothing participates / empty cloud
    ref, got = _sor_both(ctx, xyz[:50], np.zeros(50, np.uint8), 5, 0.2, 4)
    _assert_sor(ref, got)
    assert got[2][0] == 0
    keep, md, st = zim_amd.cloud_sor(np.zeros((0, 3), np.float32), None, 5, 1.0, 0.2, 4, ctx)
    assert len(keep) == 0 and st[0] == 0
    with pytest.raises(zim_amd.PsgcError):
        zim_amd.cloud_sor(xyz, None, 65, 1.0, 0.2, 4, ctx)
    with pytest.raises(zim_amd.PsgcError):
        zim_amd.cloud_sor(xyz, None, 5, 1.0, 1e-9, 4, ctx)   # 10^27 cells
    # 1.1e9 cells along x (2.8e10 in all): past the 2^30 cells per axis
    line = np.array([[0, 0, 0], [1.1e6, 0, 0]], np.float32)
    with pytest.raises(zim_amd.PsgcError):
        zim_amd.cloud_sor(line, None, 1, 1.0, 1e-3, 4, ctx)
    with pytest.raises(AssertionError):
        oracle.cloud_sor(line, None, 1, 1.0, 1e-3)


@pytest.mark.parametrize("order,voxel", [(0, 0.4), (1, 0.4), (2, 0.4), (2, 0.6), (2, 1.0),
                                         (1, 0.18), (2, 0.18), (2, 0.25)])
@pytest.mark.parametrize("sem", [0, 1])
def test_cloud_mls_random(ctx, order, voxel, sem):
    xyz, valid = _rng_cloud(5000, 13, holes=0.1)
    ref = oracle.cloud_mls(xyz, valid, 0.2, order, 0.35, 2, sem, voxel)
    got = zim_amd.cloud_mls(torch.from_numpy(xyz).cuda(), torch.from_numpy(valid).cuda(), 0.2,
                            order, 0.35, 2, sem, voxel, ctx)
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    err = np.abs(got - ref).max()
    print("cloud mls max |diff| mm", err, "bit-equal:", np.array_equal(got, ref))
    assert err <= XYZ_TOL
    assert np.array_equal(got[valid == 0], xyz[valid == 0])


@pytest.mark.parametrize("voxel", [0.45, 0.6, 1.0, 0.2, 0.3])
def test_cloud_gaussian_random(ctx, voxel):
    xyz, valid = _rng_cloud(5000, 14, holes=0.05)
    ref = oracle.cloud_gaussian(xyz, valid, 0.15, 0.4, voxel)
    got = zim_amd.cloud_gaussian(xyz, valid, 0.15, 0.4, voxel, ctx)
    err = np.abs(got - ref).max()
    print("cloud gaussian max |diff| mm", err, "bit-equal:", np.array_equal(got, ref))
    assert err <= XYZ_TOL


def _merged_views(n_views=3, rows=96, cols=128):
    """Several views of one C2 scene (the same surface, independent noise),
    reconstructed on the GPU without filters and merged into one list."""
    dcfg = scenes.c2_decode_config()
    rcfg = scenes.c2_recon_config(use_mls=0, outlier_mode=0)
    ctx = zim_amd.Context(0)
    pts = []
    for v in range(n_views):
        p = zim_amd.synth.render_c2(7, device="cuda", rows=rows, cols=cols, f=2400.0,
                                     baseline=20.0, noise_seed=100 + v)
        L = zim_amd.decode(dcfg, p.left, ctx)
        R = zim_amd.decode(dcfg, p.right, ctx)
        xyz, P, _, _ = zim_amd.reconstruct(rcfg, L[0], R[0], L[1], R[1], 70, p.Q, p.R1, ctx)
        torch.cuda.synchronize()
        x, m = xyz.cpu().numpy().reshape(-1, 3), P.cpu().numpy().reshape(-1) != 0
        pts.append(x[m])
    return np.concatenate(pts).astype(np.float32)


def test_radius_filters_reject_cells_below_half_the_radius(ctx):
    xyz, valid = _rng_cloud(500, 16)
    zim_amd.cloud_gaussian(xyz, valid, 0.15, 0.4, 0.2, ctx)        # 2 x 0.2 >= 0.4: 5x5x5
    with pytest.raises(zim_amd.PsgcError):
        zim_amd.cloud_gaussian(xyz, valid, 0.15, 0.4, 0.19, ctx)
    d = torch.from_numpy(xyz).cuda()
    with pytest.raises(zim_amd.PsgcError):
        zim_amd.cloud_mls(d, None, 0.2, 2, 0.35, 1, 1, 0.17, ctx)


def test_merged_views_filters(ctx):
    xyz = _merged_views()
    assert len(xyz) > 6000
    ref, got = _sor_both(ctx, xyz, None, 50, 1.0, 4)
    _assert_sor(ref, got)
    for voxel in (1.0, 0.5):   # 3x3x3 cells of the radius, 5x5x5 of half of it
        ref = oracle.cloud_mls(xyz, None, 0.5, 2, 1.0, 1, 1, voxel)
        got = zim_amd.cloud_mls(torch.from_numpy(xyz).cuda(), None, 0.5, 2, 1.0, 1, 1, voxel, ctx)
        torch.cuda.synchronize()
        assert np.abs(got.cpu().numpy() - ref).max() <= XYZ_TOL, voxel
        ref = oracle.cloud_gaussian(xyz, None, 0.5, 1.0, voxel)
        got = zim_amd.cloud_gaussian(xyz, None, 0.5, 1.0, voxel, ctx)
        assert np.abs(got - ref).max() <= XYZ_TOL, voxel


def _multi_object_views(n_views=4, rows=240, cols=320):
    """A small instance of the merged-cloud workload: views of the seeded
    plane + 12 spheres and boxes scene from ring poses (the full workload's
    geometry, f scaled with the width), reconstructed on the GPU without
    filters (FP64 triangulation), moved into the world frame and merged.
    Returns the cloud and, per point, whether the ray hit the plane."""
    dcfg = scenes.c2_decode_config()
    rcfg = scenes.c2_recon_config(use_mls=0, outlier_mode=0)
    c = zim_amd.Context(0)
    pts, onplane = [], []
    for v in range(n_views):
        p, R, C, obj = zim_amd.synth.render_c5_view(7, v * 4, device="cuda", rows=rows, cols=cols,
                                                    f=2400.0 * cols / 2448.0)
        L = zim_amd.decode(dcfg, p.left, c)
        Rr = zim_amd.decode(dcfg, p.right, c)
        xyz, P, _, _ = zim_amd.reconstruct(rcfg, L[0], Rr[0], L[1], Rr[1], 70, p.Q, p.R1, c)
        m = P != 0
        pts.append(zim_amd.synth.to_world(xyz[m], R, C).cpu().numpy())
        onplane.append((obj[m] == 0).cpu().numpy())
    return np.concatenate(pts).astype(np.float32), np.concatenate(onplane)


def test_multi_object_scene_merged_filters(ctx):
    """Merged views of the multi-object scene: the poses put the plane's
    points at world z = 0 (scene truth), and SOR (k = 50), MLS (order 2) and
    the Gaussian over the merged cloud equal the oracle's. Radii scale with
    this instance's coarser point spacing (about 2 mm)."""
    xyz, onplane = _multi_object_views()
    assert len(xyz) > 100_000 and 0.3 < onplane.mean() < 0.95
    assert np.median(np.abs(xyz[onplane, 2])) < 1.0
    ref, got = _sor_both(ctx, xyz, None, 50, 4.0, 4)
    _assert_sor(ref, got)
    keep = ref[0]
    for voxel in (5.0, 2.5):
        ref = oracle.cloud_mls(xyz, keep, 2.5, 2, 5.0, 1, 1, voxel)
        got = zim_amd.cloud_mls(torch.from_numpy(xyz).cuda(), torch.from_numpy(keep).cuda(), 2.5,
                                2, 5.0, 1, 1, voxel, ctx)
        torch.cuda.synchronize()
        assert np.abs(got.cpu().numpy() - ref).max() <= XYZ_TOL, voxel
        ref = oracle.cloud_gaussian(xyz, keep, 2.5, 5.0, voxel)
        got = zim_amd.cloud_gaussian(xyz, keep, 2.5, 5.0, voxel, ctx)
        assert np.abs(got - ref).max() <= XYZ_TOL, voxel


def test_voxel_autotuner_returns_a_candidate_and_same_result(ctx):
    xyz, valid = _rng_cloud(6000, 15, holes=0.1)
    d = torch.from_numpy(xyz).cuda()
    best, times = zim_amd.cloud_tune_voxel(d, None, 20, (0.2, 0.3, 0.6), 4, ctx)
    assert best in (0.2, 0.3, 0.6)
    assert len(times) == 3 and all(t > 0 for t in times.values())
    # the result does not depend on the voxel (exact k nearest)
    a = zim_amd.cloud_sor(d, None, 20, 1.0, 0.2, 4, ctx)
    b = zim_amd.cloud_sor(d, None, 20, 1.0, best, 4, ctx)
    torch.cuda.synchronize()
    assert torch.equal(a[1], b[1]) and torch.equal(a[0], b[0])

