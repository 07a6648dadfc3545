"""Host-side checks of the Python API: inputs whose dtype, shape, device or
layout would make a kernel read out of bounds or misread memory are refused
with ValueError before any address reaches the library.

Runs without a GPU: CPU tensors stand in for device tensors (the device
predicate is patched), and the context is a stub without a library, so a
call that passes every check stops with AttributeError at its first
library call."""
import numpy as np
import pytest
import torch

import zim_amd
from tests import scenes


class _Stub:
    lib = None
    h = None

    def bind_stream(self):
        return self


@pytest.fixture
def as_device(monkeypatch):
    monkeypatch.setattr(zim_amd, "_is_dev", lambda x: isinstance(x, torch.Tensor))


def _maps(rows=6, cols=8):
    return (torch.zeros(rows, cols, dtype=torch.float64), torch.zeros(rows, cols, dtype=torch.float64),
            torch.zeros(rows, cols, dtype=torch.int32), torch.zeros(rows, cols, dtype=torch.int32))


def _reconstruct(lp, rp, lk, rk):
    cfg = scenes.c2_recon_config()
    return zim_amd.reconstruct(cfg, lp, rp, lk, rk, 70, np.eye(4), np.eye(3), _Stub())


def test_reconstruct_checks_device_maps(as_device):
    lp, rp, lk, rk = _maps()
    with pytest.raises(AttributeError):        # every check passed
        _reconstruct(lp, rp, lk, rk)
    with pytest.raises(ValueError, match="r_phase: dtype"):
        _reconstruct(lp, rp.float(), lk, rk)
    with pytest.raises(ValueError, match="r_k: shape"):
        _reconstruct(lp, rp, lk, rk[:, :4])
    with pytest.raises(ValueError, match="l_k: dtype"):
        _reconstruct(lp, rp, lk.long(), rk)
    with pytest.raises(ValueError, match="r_phase: must be contiguous"):
        _reconstruct(lp, torch.zeros(8, 6, dtype=torch.float64).t(), lk, rk)
    with pytest.raises(ValueError, match="expected a device tensor"):
        _reconstruct(lp, rp.numpy(), lk, rk)


def test_reconstruct_checks_host_maps():
    lp, rp, lk, rk = [m.numpy() for m in _maps()]
    with pytest.raises(AttributeError):
        _reconstruct(lp, rp, lk, rk)
    with pytest.raises(ValueError, match="r_k: shape"):
        _reconstruct(lp, rp, lk, rk[:5])


def test_decode_checks_frames(as_device):
    cfg = scenes.c2_decode_config()
    good = [torch.zeros(6, 8, dtype=torch.uint16) for _ in range(12)]
    with pytest.raises(AttributeError):
        zim_amd.decode(cfg, good, _Stub())
    bad_type = good[:5] + [good[5].to(torch.uint8)] + good[6:]
    with pytest.raises(ValueError, match=r"frames\[5\]"):
        zim_amd.decode(cfg, bad_type, _Stub())
    bad_shape = good[:3] + [torch.zeros(6, 7, dtype=torch.uint16)] + good[4:]
    with pytest.raises(ValueError, match=r"frames\[3\]"):
        zim_amd.decode(cfg, bad_shape, _Stub())
    strided = torch.zeros(6, 16, dtype=torch.uint16)[:, ::2]
    with pytest.raises(ValueError, match="rows must be dense"):
        zim_amd.decode(cfg, good[:11] + [strided], _Stub())
    with pytest.raises(ValueError, match="sample type"):
        zim_amd.decode(cfg, [g.int() for g in good], _Stub())
    # a common row pitch (views of wider rows) is allowed
    wide = torch.zeros(12, 6, 10, dtype=torch.uint16)[:, :, :8]
    with pytest.raises(AttributeError):
        zim_amd.decode(cfg, list(wide), _Stub())


def test_batch_checks_pairs(as_device):
    dcfg, rcfg = scenes.c2_decode_config(), scenes.c2_recon_config()
    L = [torch.zeros(6, 8, dtype=torch.uint16) for _ in range(12)]

    def pair(right):
        return dict(left=L, right=right, decode=dcfg, recon=rcfg, pha_f=70, Q=np.eye(4),
                    R1=np.eye(3))
    with pytest.raises(AttributeError):
        zim_amd.Batch([pair([x.clone() for x in L])], _Stub())
    with pytest.raises(ValueError, match="right frames differ"):
        zim_amd.Batch([pair([torch.zeros(5, 8, dtype=torch.uint16) for _ in range(12)])], _Stub())
    with pytest.raises(ValueError, match="12 left and 11 right"):
        zim_amd.Batch([pair(L[:11])], _Stub())


def test_mls_and_cloud_inputs(as_device):
    xyz = torch.zeros(6, 8, 3, dtype=torch.float32)
    P = torch.zeros(6, 8, dtype=torch.uint8)
    with pytest.raises(AttributeError):
        zim_amd.applyMLSFilterToPointCloud(xyz, P, 3, 0.5, ctx=_Stub())
    with pytest.raises(ValueError, match="src_xyz: dtype"):
        zim_amd.applyMLSFilterToPointCloud(xyz.double(), P, 3, 0.5, ctx=_Stub())
    with pytest.raises(ValueError, match="P: dtype"):
        zim_amd.applyMLSFilterToPointCloud(xyz, P.bool(), 3, 0.5, ctx=_Stub())
    with pytest.raises(ValueError, match="valid: shape"):
        zim_amd.cloud_mls(xyz.reshape(-1, 3), torch.zeros(47, dtype=torch.uint8), ctx=_Stub())


def test_plane_objectives_reject_host_or_mismatched_points():
    """The plane objectives take f64 device points with one selection byte per
    point; anything else is rejected before the library sees an address."""
    import torch

    from zim_amd import autograd as zag
    x = torch.zeros((5, 3), dtype=torch.float64)
    with pytest.raises(ValueError):
        zag._plane_table([x], [torch.ones(5, dtype=torch.uint8)])
    with pytest.raises(ValueError):
        zag._plane_table([x.float()], [torch.ones(5, dtype=torch.uint8)])
