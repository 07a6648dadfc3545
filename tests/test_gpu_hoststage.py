"""Host-pointer calls (mem = PSGC_MEM_HOST) go through the context's pinned
staging chunks: pitched host frames are packed row by row, maps larger than
one chunk move in several, and the results equal those of the same calls on
device buffers bit for bit."""
import ctypes as C

import numpy as np
import pytest
import torch

import zim_amd
from zim_amd import _capi, synth
from tests import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,cols,pad", [(300, 410, 37), (1200, 1500, 64)])
def test_pitched_host_frames_decode_as_device_frames(ctx, rows, cols, pad):
    g = np.random.default_rng(rows)
    cfg = scenes.c2_decode_config()
    n = 12
    dense = [g.integers(300, 2000, (rows, cols)).astype(np.uint16) for _ in range(n)]
    # the same samples in rows of cols + pad entries (the pad holds noise)
    pitched = [np.ascontiguousarray(np.concatenate(
        [d, g.integers(0, 4096, (rows, pad)).astype(np.uint16)], axis=1)) for d in dense]
    ph = np.empty((rows, cols)); k = np.empty((rows, cols), np.int32)
    mod = np.empty((rows, cols)); val = np.empty((rows, cols), np.uint8)
    arr = (C.c_void_p * n)(*[f.ctypes.data for f in pitched])
    _capi.check(ctx.lib.psgc_decode(ctx.h, C.byref(cfg), arr, n, _capi.SAMPLE_U16, rows, cols,
                                    (cols + pad) * 2, _capi.ptr(ph), _capi.ptr(k),
                                    _capi.ptr(mod), _capi.ptr(val), _capi.MEM_HOST))
    ref = zim_amd.decode(cfg, [torch.from_numpy(f).cuda() for f in dense], ctx)
    torch.cuda.synchronize()
    for a, b in zip((ph, k, mod, val), ref):
        assert np.array_equal(a, b.cpu().numpy())


def test_full_size_host_maps_reconstruct_as_device_maps(ctx):
    """A 2448 x 2048 pair: every f64 map (40 MB) spans two staging chunks."""
    p = synth.render_c2(1000, device="cuda")
    cfg = scenes.c2_decode_config()
    L = zim_amd.decode(cfg, p.left, ctx)
    R = zim_amd.decode(cfg, p.right, ctx)
    rc = scenes.c2_recon_config(use_mls=0)
    dev_out = zim_amd.reconstruct(rc, L[0], R[0], L[1], R[1], 70, p.Q, p.R1, ctx)
    torch.cuda.synchronize()
    hL = [t.cpu().numpy() for t in L[:2]]
    hR = [t.cpu().numpy() for t in R[:2]]
    host_out = zim_amd.reconstruct(rc, hL[0], hR[0], hL[1], hR[1], 70, p.Q, p.R1, ctx)
    for a, b in zip(host_out[:2], dev_out[:2]):
        assert np.array_equal(a, b.cpu().numpy())
    # and the host-mode decode of the full-size frames
    hL2 = zim_amd.decode(cfg, [f.cpu().numpy() for f in p.left], ctx)
    for a, b in zip(hL2, L):
        assert np.array_equal(a, b.cpu().numpy())


def test_concurrent_host_calls_on_two_contexts():
    """Two host threads, each with its own context, decode host frames at the
    same time (the calls release the GIL): the host-side copies share one
    thread pool and each context its own pinned chunks; every result equals
    the device decode of the same frames."""
    import threading
    cfg = scenes.c2_decode_config()
    g = np.random.default_rng(9)
    stacks = [[g.integers(300, 2000, (1024, 1280)).astype(np.uint16) for _ in range(12)]
              for _ in range(2)]
    ctxs = [zim_amd.Context(0), zim_amd.Context(0)]
    refs = [[t.cpu().numpy() for t in zim_amd.decode(cfg, [torch.from_numpy(f).cuda()
                                                          for f in st], ctxs[0])]
            for st in stacks]
    errors = []

    def worker(i):
        try:
            for _ in range(5):
                got = zim_amd.decode(cfg, stacks[i], ctxs[i])
                if not all(np.array_equal(a, b) for a, b in zip(got, refs[i])):
                    errors.append(i)
        except Exception as e:  # reported below
            errors.append(repr(e))

    th = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th), "a decode thread did not finish"
    assert not errors, errors
