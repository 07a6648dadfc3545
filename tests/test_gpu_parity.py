"""GPU parity: the HIP path through the C ABI against the CPU restatement
(oracle/) on the same seeded inputs, stage-isolated as the reference's own
compare harness does (JSLAB_Demo_Dual.cpp:448-513).

Bars: fringe orders, validity masks, integer match columns and outlier masks
bit-exact; phase within 1e-5 rad; disparity within 1e-3 px; XYZ within
1e-3 mm.
"""
import numpy as np
import pytest
import torch

import oracle
import zim_amd
from zim_amd import _capi
from tests import scenes

pytestmark = pytest.mark.gpu

PHASE_TOL = 1e-5   # rad
DISP_TOL = 1e-3    # px
XYZ_TOL = 1e-3     # mm


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def _decode_both(cfg, frames, ctx):
    ref = oracle.decode(cfg, frames)
    got = zim_amd.decode(cfg, [dev(f) for f in frames], ctx)
    torch.cuda.synchronize()
    return ref, [host(t) for t in got]


def _assert_decode(ref, got):
    rph, rk, rmod, rv = ref
    gph, gk, gmod, gv = got
    assert np.array_equal(gv, rv), f"valid mismatch: {(gv != rv).sum()} px"
    assert np.array_equal(gk, rk), f"k mismatch: {(gk != rk).sum()} px"
    assert np.abs(gph - rph).max() <= PHASE_TOL
    assert np.abs(gmod - rmod).max() <= 1e-12
    # the shared FP64 routines make these bit-identical in practice
    print("phase max |diff|", np.abs(gph - rph).max(), "bit-equal:", np.array_equal(gph, rph))


@pytest.mark.parametrize("seed", [3, 11])
def test_decode_c2_u16(kctx, seed):
    ctx = kctx
    _, left, right = scenes.c2_small(seed)
    cfg = scenes.c2_decode_config()
    for frames in (left, right):
        ref, got = _decode_both(cfg, frames, ctx)
        _assert_decode(ref, got)
        assert ref[3].mean() > 0.5  # the scene decodes


def test_decode_variants(kctx):
    ctx = kctx
    _, left, _ = scenes.c2_small(3)
    for kw in (dict(avg_three_phase=0), dict(post_smooth=0), dict(preblur_sigma=0.0),
               dict(modulation_threshold=0.6)):
        cfg = scenes.c2_decode_config(**kw)
        ref, got = _decode_both(cfg, left, ctx)
        _assert_decode(ref, got)


def test_decode_f64_and_u8_inputs(kctx):
    ctx = kctx
    _, left, _ = scenes.c2_small(3)
    cfg = scenes.c2_decode_config()
    f64 = [f.astype(np.float64) for f in left]
    ref, got = _decode_both(cfg, f64, ctx)
    _assert_decode(ref, got)
    u8 = [(f >> 4).astype(np.uint8) for f in left]
    ref, got = _decode_both(cfg, u8, ctx)
    _assert_decode(ref, got)


def test_decode_gray_code(kctx):
    ctx = kctx
    _, left, right = scenes.c1_small(5)
    cfg = scenes.c1_gray_config()
    for frames in (left, right):
        ref, got = _decode_both(cfg, frames, ctx)
        _assert_decode(ref, got)


def test_decode_host_mode_equals_device_mode(ctx):
    _, left, _ = scenes.c2_small(3)
    cfg = scenes.c2_decode_config()
    a = zim_amd.decode(cfg, left, ctx)
    b = zim_amd.decode(cfg, [dev(f) for f in left], ctx)
    for x, y in zip(a, b):
        assert np.array_equal(x, host(y))


def _decoded_pair(seed=3):
    p, left, right = scenes.c2_small(seed)
    cfg = scenes.c2_decode_config()
    L = oracle.decode(cfg, left)
    R = oracle.decode(cfg, right)
    return p, L, R


def test_stereo_match_stage_isolated(ctx):
    p, L, R = _decoded_pair()
    ref_d, ref_mc = oracle.stereo_match(L[0], R[0], L[1], R[1], 70, 1.0)
    rows, cols = L[0].shape
    d = torch.empty((rows, cols), dtype=torch.float64, device="cuda")
    mc = torch.empty((rows, cols), dtype=torch.int32, device="cuda")
    lib = ctx.lib
    ctx.bind_stream()
    args = [dev(L[0]), dev(R[0]), dev(L[1]), dev(R[1])]
    _capi.check(lib.psgc_stereo_match(ctx.h, *[_capi.ptr(a) for a in args], rows, cols, 70, 1.0,
                                      _capi.ptr(d), _capi.ptr(mc), _capi.MEM_DEVICE))
    torch.cuda.synchronize()
    assert np.array_equal(host(mc), ref_mc)
    assert np.abs(host(d) - ref_d).max() <= DISP_TOL
    assert (ref_mc >= 0).mean() > 0.3


def test_median_reproject(kctx):
    ctx = kctx
    g = np.random.default_rng(1)
    disp = (g.random((97, 131)) * 600).astype(np.float32)
    disp[g.random(disp.shape) < 0.2] = 0
    ref = oracle.median5(disp)
    got = np.empty_like(disp)
    _capi.check(ctx.lib.psgc_median5(ctx.h, _capi.ptr(disp), _capi.ptr(got), 97, 131,
                                     _capi.MEM_HOST))
    assert np.array_equal(got, ref)
    Q = [1, 0, 0, -65, 0, 1, 0, -48, 0, 0, 0, 2400.0, 0, 0, 1 / 150.0, 0]
    R1 = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    for use_rect in (1, 0):
        rx, rP = oracle.reproject(ref.astype(np.float64), Q, R1, 0, 1000, use_rect)
        gx = np.empty_like(rx)
        gP = np.empty_like(rP)
        _capi.check(ctx.lib.psgc_reproject(ctx.h, _capi.ptr(ref), 97, 131,
                                           _capi.ptr(np.array(Q, np.float64)),
                                           _capi.ptr(np.array(R1, np.float64)), 0.0, 1000.0,
                                           use_rect, _capi.ptr(gx), _capi.ptr(gP),
                                           _capi.MEM_HOST))
        assert np.array_equal(gP, rP)
        assert np.array_equal(gx, rx)


@pytest.mark.parametrize("sem", [0, 1])
@pytest.mark.parametrize("win,sigma", [(3, 0.7), (5, 1.0)])
def test_gaussian(ctx, sem, win, sigma):
    xyz, P = scenes.rng_points(80, 96, seed=2)
    rx, rP = oracle.gaussian(xyz, P, win, sigma, sem)
    gx, gP = np.empty_like(rx), np.empty_like(rP)
    _capi.check(ctx.lib.psgc_filter_gaussian(ctx.h, _capi.ptr(xyz), _capi.ptr(P), _capi.ptr(gx),
                                             _capi.ptr(gP), 80, 96, win, sigma, sem,
                                             _capi.MEM_HOST))
    assert np.array_equal(gP, rP)
    assert np.array_equal(gx, rx)


def test_ror(ctx):
    xyz, P = scenes.rng_points(80, 96, seed=4)
    for radius, mn, win in ((0.7, 10, 10), (0.5, 5, 5), (2.0, 3, 3)):
        rx, rP = oracle.ror(xyz, P, radius, mn, win)
        gx, gP = np.empty_like(rx), np.empty_like(rP)
        _capi.check(ctx.lib.psgc_filter_ror(ctx.h, _capi.ptr(xyz), _capi.ptr(P), _capi.ptr(gx),
                                            _capi.ptr(gP), 80, 96, radius, mn, win,
                                            _capi.MEM_HOST))
        assert np.array_equal(gP, rP)
        assert np.array_equal(gx, rx)


@pytest.mark.parametrize("k,half", [(30, 4), (20, 3), (30, 2), (8, 2), (1, 1)])
def test_sor(kctx, k, half):
    """(30, 2): a 5 x 5 window never holds 30 neighbours, every point is
    dropped."""
    ctx = kctx
    xyz, P = scenes.rng_points(90, 110, seed=6)
    rx, rP, rmd, rst = oracle.sor(xyz, P, k, 1.0, half)
    gx, gP = np.empty_like(rx), np.empty_like(rP)
    gmd = np.empty_like(rmd)
    gst = np.zeros(4)
    _capi.check(ctx.lib.psgc_filter_sor(ctx.h, _capi.ptr(xyz), _capi.ptr(P), _capi.ptr(gx),
                                        _capi.ptr(gP), 90, 110, k, 1.0, half, _capi.ptr(gmd),
                                        _capi.ptr(gst), _capi.MEM_HOST))
    assert np.array_equal(gst, rst), (gst, rst)
    assert np.array_equal(gmd, rmd)
    assert np.array_equal(gP, rP)
    assert np.array_equal(gx, rx)


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("sem", [0, 1])
def test_mls(kctx, order, sem):
    ctx = kctx
    xyz, P = scenes.rng_points(64, 80, seed=8, holes=0.3)
    ref = oracle.mls(xyz, P, 4, 0.5, order, 1.5, 2, sem)
    got = np.empty_like(ref)
    _capi.check(ctx.lib.psgc_filter_mls(ctx.h, _capi.ptr(xyz), _capi.ptr(P), _capi.ptr(got),
                                        64, 80, 4, 0.5, order, 1.5, 2, sem, _capi.MEM_HOST))
    err = np.abs(got - ref).max()
    print("mls max |diff| mm", err)
    assert err <= XYZ_TOL


@pytest.mark.parametrize("sem", [0, 1])
def test_mls_non_finite_points(kctx, sem):
    """Valid points with a NaN or infinite coordinate: the reference's radius
    test (d^2 > r^2 rejects) keeps a NaN distance, so every point that has
    such a neighbour comes out NaN, and infinite distances are rejected;
    both kernel sets give the oracle's result (NaN where it has NaN)."""
    ctx = kctx
    xyz, P = scenes.rng_points(48, 64, seed=9, holes=0.2)
    P[10, 10] = P[30, 40] = P[20, 5] = P[40, 60] = 1
    xyz[10, 10, 0] = np.nan          # one coordinate NaN
    xyz[30, 40] = np.nan             # every coordinate NaN
    xyz[20, 5, 2] = np.inf           # infinite: never a neighbour
    xyz[40, 60, 1] = -np.inf
    for hw in (3, 9):
        ref = oracle.mls(xyz, P, hw, 0.5, 2, 1.5, 1, sem)
        got = np.empty_like(ref)
        _capi.check(ctx.lib.psgc_filter_mls(ctx.h, _capi.ptr(xyz), _capi.ptr(P), _capi.ptr(got),
                                            48, 64, hw, 0.5, 2, 1.5, 1, sem, _capi.MEM_HOST))
        nan_ref = np.isnan(ref).any(-1)
        assert 10 < nan_ref.sum() < 0.2 * P.size       # NaN spreads to the neighbourhoods
        assert np.array_equal(np.isnan(got), np.isnan(ref)), (hw, (np.isnan(got) != np.isnan(ref)).sum())
        fin = ~np.isnan(ref)
        assert np.array_equal(got[fin], ref[fin]), hw


@pytest.mark.parametrize("variant", ["cpu_ref", "gpu_ref", "c2"])
def test_reconstruct_chain(kctx, variant):
    ctx = kctx
    p, L, R = _decoded_pair(11)
    if variant == "cpu_ref":
        cfg = zim_amd.ReconConfig.default(radius_search=1.0, radius_min_neighbors=6)
    elif variant == "gpu_ref":
        cfg = zim_amd.ReconConfig.default_gpu(radius_search=1.0, mls_window_size=5)
    else:
        cfg = scenes.c2_recon_config()
    rx, rP, rd, rmc = oracle.reconstruct(cfg, L[0], R[0], L[1], R[1], 70, p.Q, p.R1)
    gx, gP, gd, gmc = zim_amd.reconstruct(cfg, dev(L[0]), dev(R[0]), dev(L[1]), dev(R[1]), 70,
                                          p.Q, p.R1, ctx)
    torch.cuda.synchronize()
    gx, gP, gd, gmc = host(gx), host(gP), host(gd), host(gmc)
    assert np.array_equal(gmc, rmc)
    assert np.abs(gd - rd).max() <= DISP_TOL
    assert np.array_equal(gP, rP), f"mask mismatch {(gP != rP).sum()}"
    assert np.abs(gx - rx).max() <= XYZ_TOL
    assert rP.mean() > 0.2


def test_error_paths(ctx):
    cfg = scenes.c2_decode_config(freq=(70, 0, 59))
    _, left, _ = scenes.c2_small(3)
    with pytest.raises(zim_amd.PsgcError) as e:
        zim_amd.decode(cfg, left, ctx)
    assert e.value.code == _capi.PSGC_E_INVALID
    cfg = scenes.c2_decode_config()
    with pytest.raises(zim_amd.PsgcError):
        zim_amd.decode(cfg, left[:-1], ctx)


def test_tiny_and_degenerate_images(kctx):
    ctx = kctx
    cfg = scenes.c2_decode_config()
    for rows, cols in ((1, 1), (1, 7), (5, 1), (3, 3)):
        g = np.random.default_rng(rows * 100 + cols)
        frames = [g.integers(0, 4096, (rows, cols)).astype(np.uint16) for _ in range(12)]
        ref = oracle.decode(cfg, frames)
        got = zim_amd.decode(cfg, frames, ctx)
        for r_, g_ in zip(ref, got):
            assert np.array_equal(r_, g_)
    # all-constant frames: nothing valid, nothing matched
    flat = [np.full((16, 16), 1000, np.uint16) for _ in range(12)]
    ph, k, mod, v = zim_amd.decode(cfg, flat, ctx)
    assert v.sum() == 0 and k.sum() == 0


def _pair_inputs(seed, rows, cols):
    p, left, right = scenes.c2_small(seed, rows=rows, cols=cols)
    return p, [dev(f) for f in left], [dev(f) for f in right], left, right


@pytest.mark.parametrize("graph", [False, True])
def test_batch_uneven_pairs_match_oracle(ctx, graph):
    """Two pairs of different sizes (and a CPU-semantics ROR pair) in one
    batch; each must equal the CPU restatement run on that pair alone."""
    dcfg = scenes.c2_decode_config()
    specs = [(21, 96, 128, scenes.c2_recon_config(mls_window_size=4)),
             (22, 80, 200, zim_amd.ReconConfig.default(radius_search=1.0, radius_min_neighbors=6)),
             (23, 64, 96, scenes.c2_recon_config(use_mls=0))]
    pairs, refs = [], []
    for seed, rows, cols, rcfg in specs:
        p, L, R, lh, rh = _pair_inputs(seed, rows, cols)
        pairs.append(dict(left=L, right=R, decode=dcfg, recon=rcfg, pha_f=70, Q=p.Q, R1=p.R1,
                          disp=True))
        Lr, Rr = oracle.decode(dcfg, lh), oracle.decode(dcfg, rh)
        refs.append(oracle.reconstruct(rcfg, Lr[0], Rr[0], Lr[1], Rr[1], 70, p.Q, p.R1))
    b = zim_amd.Batch(pairs, ctx, graph=graph)
    for rep in range(2):   # replays must be identical
        outs = b.run()
        torch.cuda.synchronize()
        for out, (rx, rP, rd, _) in zip(outs, refs):
            assert np.array_equal(host(out["P"]), rP), f"mask mismatch {(host(out['P']) != rP).sum()}"
            assert np.abs(host(out["xyz"]) - rx).max() <= XYZ_TOL
            assert np.abs(host(out["disp"]) - rd).max() <= DISP_TOL
    b.close()


def test_batch_autotune_keeps_results(ctx):
    dcfg = scenes.c2_decode_config()
    p, L, R, lh, rh = _pair_inputs(31, 96, 128)
    rcfg = scenes.c2_recon_config(mls_window_size=4)
    b = zim_amd.Batch([dict(left=L, right=R, decode=dcfg, recon=rcfg, pha_f=70, Q=p.Q, R1=p.R1)],
                      ctx, graph=True, autotune=True)
    out = b.run()[0]
    torch.cuda.synchronize()
    Lr, Rr = oracle.decode(dcfg, lh), oracle.decode(dcfg, rh)
    rx, rP, _, _ = oracle.reconstruct(rcfg, Lr[0], Rr[0], Lr[1], Rr[1], 70, p.Q, p.R1)
    assert np.array_equal(host(out["P"]), rP)
    assert np.abs(host(out["xyz"]) - rx).max() <= XYZ_TOL
    assert b.grid("mls") > 0 and b.grid("decode") > 0
    assert b.grid("decode_rows") in (2, 4)
    # pinned shapes (the other decode tile height, odd grid sizes) re-capture
    # the graph and leave the results unchanged
    rows = 6 - b.grid("decode_rows")
    variant = 1 - b.grid("mls_variant")
    mt = 128 if b.grid("match_threads") != 128 else 512
    for stage, v in (("decode_rows", rows), ("mls_variant", variant), ("match_threads", mt),
                     ("decode", 37), ("match", 300), ("mls", 257), ("sor_knn", 1)):
        b.set_grid(stage, v)
        assert b.grid(stage) == v
    for run in (b.run, b.run_eager):
        out["xyz"].fill_(float("nan"))
        out["P"].fill_(7)
        run()
        torch.cuda.synchronize()
        assert np.array_equal(host(out["P"]), rP)
        assert np.abs(host(out["xyz"]) - rx).max() <= XYZ_TOL
    for stage, v in (("decode_rows", 3), ("mls_variant", 2), ("match_threads", 64), ("decode", 0),
                     ("nope", 64)):
        with pytest.raises(zim_amd.PsgcError):
            b.set_grid(stage, v)
    b.close()


def _inradius_counts(xyz, P, hw, r):
    """In-radius valid neighbours (the point itself included) of every valid
    point over the (2 hw + 1)^2 window, with the MLS filter's float32
    squared distance."""
    rows, cols = P.shape
    x = xyz.astype(np.float32)
    ok = P != 0
    cnt = np.zeros((rows, cols), np.int32)
    for dy in range(-hw, hw + 1):
        for dx in range(-hw, hw + 1):
            ys, yd = slice(max(dy, 0), rows + min(dy, 0)), slice(max(-dy, 0), rows + min(-dy, 0))
            xs, xd = slice(max(dx, 0), cols + min(dx, 0)), slice(max(-dx, 0), cols + min(-dx, 0))
            d = x[yd, xd] - x[ys, xs]
            d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]
            cnt[yd, xd] += (ok[yd, xd] & ok[ys, xs] & (d2 <= np.float32(r * r))).astype(np.int32)
    return np.where(ok, cnt, 0)


def test_batch_mls_variants_match_oracle(ctx):
    """Both MLS kernels, pinned in turn on one captured batch: weights
    evaluated in every pass, and weights kept in LDS between the passes
    (points with more in-radius neighbours than the cache holds evaluate the
    rest again); each against the CPU restatement. The batch holds the pair
    six times, with half windows 6, 8, 10, 12 and 14 (one launch per window
    class, caches of 24, 20 and 16 weights, none above 10) under GPU
    semantics (fused sums) and CPU semantics (rounded products and sums)."""
    dcfg = scenes.c2_decode_config()
    # a 2.5 mm radius on this ~1 m deep scene gives many points more
    # neighbours than any cache holds
    p, L, R, lh, rh = _pair_inputs(41, 128, 192)
    Lr, Rr = oracle.decode(dcfg, lh), oracle.decode(dcfg, rh)
    nx, nP, _, _ = oracle.reconstruct(scenes.c2_recon_config(use_mls=0), Lr[0], Rr[0], Lr[1],
                                      Rr[1], 70, p.Q, p.R1)
    specs = [(6, _capi.SEM_GPU, 24), (6, _capi.SEM_CPU, 24), (8, _capi.SEM_GPU, 20),
             (10, _capi.SEM_CPU, 16), (12, _capi.SEM_GPU, 0), (14, _capi.SEM_CPU, 0)]
    cfgs, refs = [], []
    for hw, sem, kc in specs:
        cnt = _inradius_counts(nx, nP, hw, 2.5)
        print(f"hw {hw}: in-radius neighbours max {cnt.max()}, points above {kc}: "
              f"{int((cnt > kc).sum())}", flush=True)
        if kc:
            assert (cnt > kc).sum() > 100 and ((cnt > 3) & (cnt <= kc)).sum() > 100
        c = scenes.c2_recon_config(mls_max_radius=2.5, mls_spatial_sigma=1.0,
                                   mls_window_size=hw, semantics=sem)
        cfgs.append(c)
        refs.append(oracle.reconstruct(c, Lr[0], Rr[0], Lr[1], Rr[1], 70, p.Q, p.R1)[:2])
    m = (refs[0][1] != 0) & (refs[1][1] != 0)
    assert not np.array_equal(refs[0][0][m], refs[1][0][m])   # the two roundings differ
    b = zim_amd.Batch([dict(left=L, right=R, decode=dcfg, recon=c, pha_f=70, Q=p.Q, R1=p.R1)
                       for c in cfgs], ctx, graph=True)
    for variant in (1, 0, 1):
        b.set_grid("mls_variant", variant)
        outs = b.run()
        torch.cuda.synchronize()
        for out, (ex, eP) in zip(outs, refs):
            assert np.array_equal(host(out["P"]), eP)
            assert np.array_equal(host(out["xyz"]), ex), np.abs(host(out["xyz"]) - ex).max()
    b.close()


# ------------------------------------------------------------------- HDR --

@pytest.mark.parametrize("smooth", [0, 1])
def test_decode_hdr_sweep(kctx, smooth):
    """Clipped noise-free fringes (f64): the device HDR fit against the
    oracle's, including the all-saturated rows (invalid, order -1)."""
    frames, _, _ = scenes.hdr_sweep(cols=640)
    cfg = scenes.hdr_config(post_smooth=smooth)
    ref, got = _decode_both(cfg, frames, kctx)
    _assert_decode(ref, got)
    assert (got[1][6:] == -1).all()


def test_decode_hdr_saturated_u8(kctx):
    _, left, right = scenes.hdr_small()
    cfg = scenes.hdr_config()
    for frames in (left, right):
        assert np.mean([(f == 255).mean() for f in frames]) > 0.05  # really over-exposed
        ref, got = _decode_both(cfg, frames, kctx)
        _assert_decode(ref, got)
        assert ref[3].mean() > 0.5


def test_batch_mixed_schedules_match_oracle(ctx):
    """One captured batch of uneven pairs with different schedules: 3-step
    and 8-step heterodyne (uint16), the 640x480 6-bit Gray-code pair of the
    CPU configuration (uint8, CPU semantics, SOR k=20) and an over-exposed
    8-step pair decoded with the HDR fit (uint8). Each pair must
    equal the oracle run on it alone, on every replay."""
    from zim_amd import synth
    specs = []
    for seed, rows, cols, steps in ((41, 64, 96, 3), (42, 80, 120, 8)):
        p = synth.render_c2(seed, rows=rows, cols=cols, f=2400.0, baseline=20.0,
                            steps=(steps,) * 3)
        specs.append((p, [x.numpy() for x in p.left], [x.numpy() for x in p.right],
                      scenes.c2_decode_config(shift_num=(steps,) * 3),
                      scenes.c2_recon_config(mls_window_size=4, max_depth=2000.0), 70))
    p = synth.render_c1(5)   # the CPU configuration: 640x480, 6-bit Gray code, SOR k=20
    specs.append((p, [x.numpy() for x in p.left], [x.numpy() for x in p.right],
                  scenes.c1_gray_config(),
                  zim_amd.ReconConfig.default(outlier_mode=_capi.OUTLIER_SOR, sor_k=20,
                                              sor_std_mul=1.0, sor_window=4), 50))
    p, lh, rh = scenes.hdr_small()
    specs.append((p, lh, rh, scenes.hdr_config(), scenes.c2_recon_config(mls_window_size=4), 70))
    pairs, refs = [], []
    for p, lh, rh, dcfg, rcfg, pha_f in specs:
        pairs.append(dict(left=[dev(f) for f in lh], right=[dev(f) for f in rh], decode=dcfg,
                          recon=rcfg, pha_f=pha_f, Q=p.Q, R1=p.R1, disp=True))
        Lr, Rr = oracle.decode(dcfg, lh), oracle.decode(dcfg, rh)
        refs.append(oracle.reconstruct(rcfg, Lr[0], Rr[0], Lr[1], Rr[1], pha_f, p.Q, p.R1))
    b = zim_amd.Batch(pairs, ctx, graph=True)
    for rep in range(3):
        outs = b.run()
        torch.cuda.synchronize()
        for i, (out, (rx, rP, rd, _)) in enumerate(zip(outs, refs)):
            assert rP.sum() > 0, i
            assert np.array_equal(host(out["P"]), rP), (rep, i, (host(out["P"]) != rP).sum())
            assert np.abs(host(out["xyz"]) - rx).max() <= XYZ_TOL, (rep, i)
            assert np.abs(host(out["disp"]) - rd).max() <= DISP_TOL, (rep, i)
    b.close()


def test_batch_at_uneven_full_sizes_matches_oracle(ctx):
    """One captured, autotuned batch of the uneven configuration's real sizes
    and schedules, every pair the config-1 scene in uint16 12-bit: 640x480
    3-step heterodyne, 1280x1024 4-step Gray code, 1920x1200 8-step
    heterodyne and 2448x2048 3-step Gray code; each pair equal to the oracle
    run on it alone, with the row match in one launch and per image width
    (identical outputs)."""
    from zim_amd import synth
    specs = []
    for i, (rows, cols, steps, gray) in enumerate(((480, 640, 3, False), (1024, 1280, 4, True),
                                                   (1200, 1920, 8, False),
                                                   (2048, 2448, 3, True))):
        p = synth.render_c3(4100 + i, device="cuda", rows=rows, cols=cols, steps=steps,
                            gray=gray)
        if gray:
            dcfg, pha_f = scenes.c1_gray_config(shift_num=(steps, 0, 0)), 50
        else:
            dcfg, pha_f = scenes.c2_decode_config(shift_num=(steps,) * 3), 70
        specs.append((p, dcfg, pha_f))
    rcfg = scenes.c2_recon_config()
    b = zim_amd.Batch([dict(left=p.left, right=p.right, decode=d, recon=rcfg, pha_f=f, Q=p.Q,
                            R1=p.R1) for p, d, f in specs], ctx, graph=True, autotune=True)
    for w in (640, 1280, 1920, 2448):
        assert b.grid(f"match_threads@{w}") in (128, 256, 512) and b.grid(f"match@{w}") > 0
    chosen = b.grid("match_groups")
    assert chosen in (0, 4)
    outs = [{k: v.clone() for k, v in o.items()} for o in b.run()]
    b.set_grid("match_groups", 0 if chosen else 1)
    assert b.grid("match_groups") == (0 if chosen else 4)
    again = b.run()
    torch.cuda.synchronize()
    for o, a in zip(outs, again):
        assert torch.equal(o["P"], a["P"]) and torch.equal(o["xyz"], a["xyz"])
    for i, ((p, dcfg, pha_f), out) in enumerate(zip(specs, outs)):
        lh = [x.cpu().numpy() for x in p.left]
        rh = [x.cpu().numpy() for x in p.right]
        Lr, Rr = oracle.decode(dcfg, lh), oracle.decode(dcfg, rh)
        rx, rP, _, _ = oracle.reconstruct(rcfg, Lr[0], Rr[0], Lr[1], Rr[1], pha_f, p.Q, p.R1)
        print("pair", i, p.left[0].shape, "valid", int(rP.sum()), flush=True)
        assert rP.sum() > 0.3 * rP.size, i
        assert np.array_equal(host(out["P"]), rP), (i, (host(out["P"]) != rP).sum())
        assert np.abs(host(out["xyz"]) - rx).max() <= XYZ_TOL, i
    with pytest.raises(zim_amd.PsgcError):
        b.grid("match@1000")
    b.close()


@pytest.mark.parametrize("seed", [1000, 1031])
def test_full_size_pair_matches_oracle(ctx, seed):
    """A full 2448x2048 config-2 pair (rendered on the GPU, decoded and
    reconstructed by the batched pipeline) against the oracle on the same
    frames: masks bit-exact, XYZ within 1e-3 mm over millions of points. The
    MLS window is the pair's own (seed 1031 is the nearest pair of the
    benchmark, half window 10), and no in-radius neighbour of the MLS input
    lies on its outer ring: the window holds the 1.5 mm radius."""
    from zim_amd import synth
    p = synth.render_c2(seed, device="cuda")
    dcfg = scenes.c2_decode_config()
    hw = scenes.c2_window(p)
    rcfg = scenes.c2_recon_config(mls_window_size=hw)
    print(f"[full-size] seed {seed}: MLS half window {hw}", flush=True)
    b = zim_amd.Batch([dict(left=p.left, right=p.right, decode=dcfg, recon=rcfg, pha_f=70,
                            Q=p.Q, R1=p.R1)], ctx, graph=False)
    out = b.run()[0]
    torch.cuda.synchronize()
    lh = [host(f) for f in p.left]
    rh = [host(f) for f in p.right]
    Lr, Rr = oracle.decode(dcfg, lh), oracle.decode(dcfg, rh)
    rx, rP, _, _ = oracle.reconstruct(rcfg, Lr[0], Rr[0], Lr[1], Rr[1], 70, p.Q, p.R1)
    gP, gx = host(out["P"]), host(out["xyz"])
    m = rP != 0
    print("points", int(m.sum()), "max |dxyz|", np.abs(gx[m] - rx[m]).max())
    assert m.sum() > 2_000_000
    sx, sP, _, _ = oracle.reconstruct(scenes.c2_recon_config(use_mls=0), Lr[0], Rr[0], Lr[1],
                                      Rr[1], 70, p.Q, p.R1)
    assert scenes.ring_points(sx, sP, hw, 1.5) == 0
    assert np.array_equal(gP, rP), f"mask mismatch {(gP != rP).sum()}"
    assert np.abs(gx - rx).max() <= XYZ_TOL
    b.close()


# ----------------------------------------------- polynomial triangulation --

def _mixed_poly_maps(rows, cols, seed=2):
    """Coefficient maps whose rows cycle through every form (the three
    Q-equivalent ones and three polynomial ones) and the masking cases."""
    f, cx1, cy, cx2, B = 1600.0, cols / 2 + 0.5, rows / 2 - 0.5, cols / 2 - 4.0, 120.0
    forms = [scenes.poly_coefs_equivalent_to_q(rows, cols, f, cx1, cy, cx2, B, fm)
             for fm in ("inv_cubic", "rational", "inv_quad")]
    g = np.random.default_rng(seed)
    cx = g.uniform(-1, 1, (rows, cols, 4)).astype(np.float32) * np.float32([1, 1e-2, 1e-4, 1e-6])
    cy = g.uniform(-1, 1, (rows, cols, 4)).astype(np.float32) * np.float32([1, 1e-2, 1e-4, 1e-6])
    cz = np.stack([np.full((rows, cols), 700.0), g.uniform(-2, 2, (rows, cols)),
                   g.uniform(-1e-2, 1e-2, (rows, cols)), g.uniform(-1e-4, 1e-4, (rows, cols))],
                  -1).astype(np.float32)
    out = [cx.copy(), cy.copy(), cz.copy()]
    for r in range(rows):
        k = r % 7
        if k < 3:
            for a in range(3):
                out[a][r] = forms[k][a][r]
        elif k == 3:
            out[0][r, :, 3] = -2.0
        elif k == 4:
            out[0][r, :, 3] = -3.0
        elif k == 5:
            out[2][r, ::3] = 0.0
    return out


def test_reproject_poly_stage(ctx):
    rows, cols = 70, 90
    g = np.random.default_rng(3)
    d = g.uniform(20.0, 90.0, (rows, cols))
    d[g.random((rows, cols)) < 0.1] = 0.0
    coef = _mixed_poly_maps(rows, cols)
    a = np.radians(2.0)
    R1 = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    for use_rect in (1, 0):
        rx, rP = oracle.reproject_poly(d, coef, R1, 0.0, 2e4, use_rect)
        gx, gP = zim_amd.reproject_poly(dev(d), [dev(c) for c in coef], R1, 0.0, 2e4, use_rect,
                                        ctx=ctx)
        torch.cuda.synchronize()
        assert rP.sum() > 0.5 * rP.size
        assert np.array_equal(host(gP), rP)
        assert np.array_equal(host(gx), rx)     # same double arithmetic: bit-identical


@pytest.mark.parametrize("variant", ["gpu_ref", "c2"])
def test_reconstruct_poly_chain(kctx, variant):
    """startCalculate_v2 on device (both kernel sets) against the oracle."""
    p, L, R = _decoded_pair(3)
    cam = p.cam
    rows, cols = L[0].shape
    coef = scenes.poly_coefs_equivalent_to_q(rows, cols, cam.f, cam.cx, cam.cy, cam.cx,
                                             cam.baseline, "inv_cubic")
    # the reference GPU defaults with a radius that fits the 0.3 mm pixel
    # pitch and an MLS window the tiled kernels hold (as in the Q chain test)
    rc = (zim_amd.ReconConfig.default_gpu(radius_search=1.0, mls_window_size=5)
          if variant == "gpu_ref" else scenes.c2_recon_config(mls_window_size=4))
    rx, rP, rd, rmc = oracle.reconstruct_poly(rc, L[0], R[0], L[1], R[1], 70, coef, p.R1)
    gx, gP, gd, gmc = zim_amd.reconstruct_poly(
        rc, *[dev(a) for a in (L[0], R[0], L[1], R[1])], 70, [dev(c) for c in coef], p.R1,
        ctx=kctx)
    torch.cuda.synchronize()
    assert rP.sum() > 1000
    assert np.array_equal(host(gmc), rmc)
    assert np.array_equal(host(gP), rP)
    assert np.abs(host(gx) - rx).max() <= XYZ_TOL
    assert np.abs(host(gd) - rd).max() <= DISP_TOL


def test_batch_with_poly_pair(ctx):
    """A batch mixing a Q pair and a polynomial-triangulation pair."""
    dcfg = scenes.c2_decode_config()
    rcfg = scenes.c2_recon_config(mls_window_size=4)
    pairs, refs = [], []
    for k, seed in enumerate((51, 52)):
        p, Ld, Rd, lh, rh = _pair_inputs(seed, 96, 128)
        Lr, Rr = oracle.decode(dcfg, lh), oracle.decode(dcfg, rh)
        spec = dict(left=Ld, right=Rd, decode=dcfg, recon=rcfg, pha_f=70, Q=p.Q, R1=p.R1)
        if k == 1:
            cam = p.cam
            coef = scenes.poly_coefs_equivalent_to_q(96, 128, cam.f, cam.cx, cam.cy, cam.cx,
                                                     cam.baseline, "inv_quad")
            spec["coef"] = [dev(c) for c in coef]
            refs.append(oracle.reconstruct_poly(rcfg, Lr[0], Rr[0], Lr[1], Rr[1], 70, coef,
                                                p.R1))
        else:
            refs.append(oracle.reconstruct(rcfg, Lr[0], Rr[0], Lr[1], Rr[1], 70, p.Q, p.R1))
        pairs.append(spec)
    b = zim_amd.Batch(pairs, ctx, graph=True)
    for rep in range(2):
        outs = b.run()
        torch.cuda.synchronize()
        for out, (rx, rP, _, _) in zip(outs, refs):
            assert rP.sum() > 0
            assert np.array_equal(host(out["P"]), rP)
            assert np.abs(host(out["xyz"]) - rx).max() <= XYZ_TOL
    b.close()


# ------------------------------------------------ ROI-restricted reconstruction --

@pytest.mark.parametrize("variant", ["ror_mls", "sor", "ror_fallback"])
def test_reconstruct_poly_roi(ctx, variant):
    """startCalculate_v2_roi on device against the oracle, with the right
    window from the device (checked against the oracle's)."""
    p, L, R = _decoded_pair(3)
    cam = p.cam
    rows, cols = L[0].shape
    coef = scenes.poly_coefs_equivalent_to_q(rows, cols, cam.f, cam.cx, cam.cy, cam.cx,
                                             cam.baseline, "inv_cubic")
    lr = [30, 17, 180, 150]
    g = np.random.default_rng(4)
    mask = np.ones((150, 180), np.uint8)
    mask[g.random((150, 180)) < 0.05] = 0
    rr_ref = oracle.roi_right_window(L[1], R[1], 70, lr, mask)
    rr = zim_amd.roi_right_window(dev(L[1]), dev(R[1]), 70, lr, dev(mask), ctx=ctx)
    assert rr.tolist() == rr_ref.tolist()
    if variant == "ror_mls":
        rc = zim_amd.ReconConfig.default(radius_search=1.0, radius_min_neighbors=6, use_mls=1,
                                         mls_window_size=4, mls_iterations=2)
    elif variant == "sor":
        rc = scenes.c2_recon_config(mls_window_size=4)
    else:
        rc = zim_amd.ReconConfig.default(radius_search=1e-4, radius_min_neighbors=8)
    rx, rP, rd = oracle.reconstruct_poly_roi(rc, L[0], R[0], L[1], R[1], 70, lr, rr, mask, coef,
                                             p.R1)
    gx, gP, gd = zim_amd.reconstruct_poly_roi(rc, *[dev(a) for a in (L[0], R[0], L[1], R[1])], 70,
                                              lr, rr, dev(mask), [dev(c) for c in coef], p.R1,
                                              ctx=ctx)
    torch.cuda.synchronize()
    assert rP.sum() > 1000
    assert np.array_equal(host(gd), rd)
    assert np.array_equal(host(gP), rP)
    assert np.abs(host(gx) - rx).max() <= XYZ_TOL
