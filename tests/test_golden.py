"""Committed golden fixtures (tests/golden/make_golden.py): the oracle must
keep reproducing them bit for bit, and the HIP path must match them to the
parity bars without running the oracle."""
import os

import numpy as np
import pytest

import oracle
from tests import scenes

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


def test_oracle_reproduces_c2_fixture():
    z = load("c2_tiny.npz")
    dc = scenes.c2_decode_config()
    L = oracle.decode(dc, list(z["left"]))
    R = oracle.decode(dc, list(z["right"]))
    for got, key in zip(L, ("l_phase", "l_k", "l_mod", "l_valid")):
        assert np.array_equal(got, z[key])
    assert np.array_equal(R[0], z["r_phase"]) and np.array_equal(R[1], z["r_k"])
    rc = scenes.c2_recon_config(mls_window_size=3)
    xyz, P, disp, mc = oracle.reconstruct(rc, L[0], R[0], L[1], R[1], 70, z["Q"], z["R1"])
    assert np.array_equal(P, z["P"]) and np.array_equal(mc, z["match_col"])
    assert np.array_equal(xyz, z["xyz"]) and np.array_equal(disp, z["disp"])


def test_oracle_reproduces_c1_fixture():
    z = load("c1_tiny.npz")
    g = oracle.decode(scenes.c1_gray_config(), list(z["left"]))
    for got, key in zip(g, ("phase", "k", "mod", "valid")):
        assert np.array_equal(got, z[key])


@pytest.mark.gpu
@pytest.mark.parametrize("kernels", ["tiled", "plain"])
def test_hip_path_matches_fixtures(kernels):
    import torch
    import zim_amd
    ctx = zim_amd.Context(0).set_kernels(kernels)
    z = load("c2_tiny.npz")
    dc = scenes.c2_decode_config()
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    L = zim_amd.decode(dc, [dev(f) for f in z["left"]], ctx)
    R = zim_amd.decode(dc, [dev(f) for f in z["right"]], ctx)
    torch.cuda.synchronize()
    assert np.array_equal(L[1].cpu().numpy(), z["l_k"])
    assert np.array_equal(L[3].cpu().numpy(), z["l_valid"])
    assert np.abs(L[0].cpu().numpy() - z["l_phase"]).max() <= 1e-5
    rc = scenes.c2_recon_config(mls_window_size=3)
    xyz, P, disp, mc = zim_amd.reconstruct(rc, L[0], R[0], L[1], R[1], 70, z["Q"], z["R1"], ctx)
    torch.cuda.synchronize()
    assert np.array_equal(P.cpu().numpy(), z["P"])
    assert np.array_equal(mc.cpu().numpy(), z["match_col"])
    assert np.abs(xyz.cpu().numpy() - z["xyz"]).max() <= 1e-3
    g1 = load("c1_tiny.npz")
    G = zim_amd.decode(scenes.c1_gray_config(), [dev(f) for f in g1["left"]], ctx)
    torch.cuda.synchronize()
    assert np.array_equal(G[1].cpu().numpy(), g1["k"])
    assert np.array_equal(G[3].cpu().numpy(), g1["valid"])
