"""Configurations beyond what the tiled kernels hold run the pair's stage on
the one-thread-per-pixel kernels, inside the same plan (and captured graph):
the reference GPU class's own defaults (MLS half window 25,
JSLAB_PhaseEpipolarReconstructor_V2_GPU.h:54-60), SOR windows above half
size 8, and rows too wide for the match's LDS row. Each against the oracle."""
import numpy as np
import pytest
import torch

import oracle
import zim_amd
from tests import scenes

pytestmark = pytest.mark.gpu

XYZ_TOL = 1e-3    # mm
DISP_TOL = 1e-3   # px


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _decoded(p_l_r, dcfg):
    _, left, right = p_l_r
    return oracle.decode(dcfg, left), oracle.decode(dcfg, right)


def _check(got, ref):
    gx, gP = got[0].cpu().numpy(), got[1].cpu().numpy()
    rx, rP = ref[0], ref[1]
    assert np.array_equal(gP, rP), f"mask mismatch {(gP != rP).sum()}"
    assert np.abs(gx - rx).max() <= XYZ_TOL


def test_reconstruct_with_the_reference_gpu_defaults(kctx):
    """psgc_recon_config_default_gpu() unmodified, on a close-range pair
    where its 0.05 mm outlier radius keeps most points."""
    pl = scenes.closeup()
    p = pl[0]
    L, R = _decoded(pl, scenes.c2_decode_config())
    rc = zim_amd.ReconConfig.default_gpu()
    assert rc.mls_window_size == 25 and rc.use_mls == 1
    print("[fallback] default_gpu on", L[0].shape, flush=True)
    ref = oracle.reconstruct(rc, L[0], R[0], L[1], R[1], 70, p.Q, p.R1)
    got = zim_amd.reconstruct(rc, dev(L[0]), dev(R[0]), dev(L[1]), dev(R[1]), 70, p.Q, p.R1,
                              kctx)
    torch.cuda.synchronize()
    _check(got, ref)
    assert ref[1].sum() > 10000
    assert np.array_equal(got[3].cpu().numpy(), ref[3])


@pytest.mark.parametrize("graph", [False, True])
def test_batch_mixes_tiled_and_fallback_pairs(ctx, graph):
    """One batch: the reference GPU defaults (plain MLS after ROR), a pair
    with an SOR window of half size 12 and a tiled MLS, a pair with a tiled
    SOR and a half-25 MLS (the SOR applied before the plain MLS), and a tiled
    pair; two replays, each pair equal to the oracle."""
    dcfg = scenes.c2_decode_config()
    cp, cl, cr = scenes.closeup()
    specs = [((cp, cl, cr), zim_amd.ReconConfig.default_gpu()),
             (scenes.c2_small(51, rows=96, cols=128), scenes.c2_recon_config(sor_window=12,
                                                                              mls_window_size=5)),
             (scenes.c2_small(52, rows=80, cols=112), scenes.c2_recon_config(sor_window=4,
                                                                              mls_window_size=10)),
             (scenes.c2_small(53, rows=64, cols=96), scenes.c2_recon_config(mls_window_size=4))]
    pairs, refs = [], []
    for (p, left, right), rc in specs:
        pairs.append(dict(left=[dev(f) for f in left], right=[dev(f) for f in right],
                          decode=dcfg, recon=rc, pha_f=70, Q=p.Q, R1=p.R1, disp=True))
        Lr, Rr = oracle.decode(dcfg, left), oracle.decode(dcfg, right)
        refs.append(oracle.reconstruct(rc, Lr[0], Rr[0], Lr[1], Rr[1], 70, p.Q, p.R1))
    b = zim_amd.Batch(pairs, ctx, graph=graph)
    for rep in range(2):
        print(f"[fallback] mixed batch replay {rep}", flush=True)
        outs = b.run()
        torch.cuda.synchronize()
        for out, ref in zip(outs, refs):
            _check((out["xyz"], out["P"]), ref)
            assert np.abs(out["disp"].cpu().numpy() - ref[2]).max() <= DISP_TOL
    b.close()


@pytest.mark.parametrize("cols", [6000, 7000])
def test_rows_wider_than_the_match_lds_row(kctx, ctx, cols):
    """Wide rows through the single-pair path and the batch (whose decode
    leaves the post-smooth to the match): 6000 px fit the 64 KB row budget
    of the tiled match (10 B per pixel), 7000 px run the one-thread-per-pixel
    match."""
    p, left, right = scenes.c2_small(61, rows=8, cols=cols)
    dcfg = scenes.c2_decode_config()
    Lr, Rr = oracle.decode(dcfg, left), oracle.decode(dcfg, right)
    rc = scenes.c2_recon_config(mls_window_size=4)
    ref = oracle.reconstruct(rc, Lr[0], Rr[0], Lr[1], Rr[1], 70, p.Q, p.R1)
    assert ref[1].sum() > 5000
    print("[fallback] wide rows, single pair", flush=True)
    got = zim_amd.reconstruct(rc, dev(Lr[0]), dev(Rr[0]), dev(Lr[1]), dev(Rr[1]), 70, p.Q, p.R1,
                              kctx)
    torch.cuda.synchronize()
    _check(got, ref)
    assert np.array_equal(got[3].cpu().numpy(), ref[3])
    print("[fallback] wide rows, batch", flush=True)
    b = zim_amd.Batch([dict(left=[dev(f) for f in left], right=[dev(f) for f in right],
                            decode=dcfg, recon=rc, pha_f=70, Q=p.Q, R1=p.R1)], ctx, graph=True)
    out = b.run()[0]
    torch.cuda.synchronize()
    _check((out["xyz"], out["P"]), ref)
    b.close()
