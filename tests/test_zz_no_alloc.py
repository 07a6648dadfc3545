"""psgc_compact_noalloc (no counterpart in the reference: compaction for
callers that cannot allocate) against psgc_compact, through the C ABI, on
two calls in a row (the second smaller), with the outputs bit for bit."""
import ctypes as C

import numpy as np
import pytest
import torch

from zim_amd import _capi

pytestmark = pytest.mark.gpu


def _inputs(rows, cols, density, seed, with_gray):
    g = np.random.default_rng(seed)
    xyz = g.normal(0, 100, (rows, cols, 3)).astype(np.float32)
    P = (g.random((rows, cols)) < density).astype(np.uint8)
    gray = g.integers(0, 256, (rows, cols)).astype(np.uint8) if with_gray else None
    return xyz, P, gray


def _dptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _run_both(ctx, xyz, P, gray):
    rows, cols = P.shape
    n_px = rows * cols
    dx = torch.from_numpy(xyz).cuda()
    dP = torch.from_numpy(P).cuda()
    dg = torch.from_numpy(gray).cuda() if gray is not None else None
    outs = []
    for fn in ("psgc_compact", "psgc_compact_noalloc"):
        ox = torch.full((n_px, 3), float("nan"), dtype=torch.float32, device="cuda")
        og = torch.full((n_px,), 7, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        if fn == "psgc_compact":
            cnt = C.c_int64(-1)
            _capi.check(ctx.lib.psgc_compact(ctx.h, _dptr(dx), _dptr(dP), _dptr(dg), rows, cols,
                                             _dptr(ox), _dptr(og), C.byref(cnt),
                                             _capi.MEM_DEVICE))
            n = cnt.value
        else:
            dcnt = torch.full((1,), -1, dtype=torch.int64, device="cuda")
            torch.cuda.synchronize()
            _capi.check(ctx.lib.psgc_compact_noalloc(ctx.h, _dptr(dx), _dptr(dP), _dptr(dg),
                                                     rows, cols, _dptr(ox), _dptr(og),
                                                     _dptr(dcnt)))
            _capi.check(ctx.lib.psgc_synchronize(ctx.h))
            n = int(dcnt.cpu()[0])
        _capi.check(ctx.lib.psgc_synchronize(ctx.h))
        outs.append((n, ox[:n].cpu().numpy(), og[:n].cpu().numpy()))
    return outs


def test_noalloc_compaction_equals_compaction_on_two_calls(ctx):
    cases = [(1024, 1280, 0.4, 1, True), (300, 401, 0.7, 2, False)]
    for rows, cols, density, seed, with_gray in cases:
        print(f"[no_alloc] compaction {rows}x{cols} density {density}", flush=True)
        xyz, P, gray = _inputs(rows, cols, density, seed, with_gray)
        (n0, x0, g0), (n1, x1, g1) = _run_both(ctx, xyz, P, gray)
        n = int(P.sum())
        assert n0 == n and n1 == n
        assert np.array_equal(x1, x0) and np.array_equal(g1, g0)
        assert np.array_equal(x1, xyz[P > 0])
        assert np.array_equal(g1, gray[P > 0] if gray is not None else np.full(n, 255, np.uint8))


def test_noalloc_compaction_empty_mask_and_limits(ctx):
    xyz, P, _ = _inputs(64, 80, 0.0, 3, False)
    (n0, _, _), (n1, _, _) = _run_both(ctx, xyz, P, None)
    assert n0 == 0 and n1 == 0
    # sizes beyond the static scratch, and null pointers, are refused on the
    # host before anything is launched
    t = torch.zeros(16, dtype=torch.float32, device="cuda")
    p = torch.zeros(16, dtype=torch.uint8, device="cuda")
    c = torch.zeros(1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    rows, cols = 1 << 14, (1 << 13) + 1       # just over 2^27 pixels
    rc = ctx.lib.psgc_compact_noalloc(ctx.h, _dptr(t), _dptr(p), None, rows, cols, _dptr(t),
                                      None, _dptr(c))
    assert rc == _capi.PSGC_E_INVALID
    assert "allocation-free limit" in ctx.lib.psgc_last_error().decode()
    rc = ctx.lib.psgc_compact_noalloc(ctx.h, None, _dptr(p), None, 4, 4, _dptr(t), None,
                                      _dptr(c))
    assert rc == _capi.PSGC_E_INVALID
